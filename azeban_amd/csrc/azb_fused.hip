// Fused row kernels and k-space kernel launches (the kernels are in
// azb_kernels.hpp).  Compiled as a separate translation unit with the
// instruction-level-parallelism scheduler, which measured faster for these
// kernels and slower for the column passes.
#include <hip/hip_runtime.h>

#include "azb_launch.hpp"

namespace azb {
namespace {

template <typename T, int L, int P>
void launch_row4_P(const RowArgs& a, long nsamples, hipStream_t st) {
  const long blocks = a.nlines / (2 * P) * nsamples;
  hipLaunchKernelGGL((euler_row4_kernel<T, L, P>), dim3((unsigned)blocks), dim3(Row4Cfg<T, L, P>::NT), 0, st, a);
}

// pdiv in {1, 2}: where the row-4 kernel runs (3D L = 96; FP32 at L = 192) it
// takes P0 = 1536/L line pairs per workgroup or half of that; every other kernel
// has one line-pair tile
template <typename T, int L, int DIM>
void launch_row_L(const RowArgs& a, long nsamples, hipStream_t st, int pdiv) {
  if constexpr (DIM == 3 && sizeof(T) == 4 && L == 768) {
    // FP32 at L = 768: the radix-24 kernel, two line pairs per workgroup, whatever
    // pdiv says (one transform plan per precision and size keeps results
    // independent of the tuning)
    const long blocks = a.nlines / (2 * Row8Cfg<T, L>::P) * nsamples;
    hipLaunchKernelGGL((euler_row8_kernel<T, L>), dim3((unsigned)blocks), dim3(Row8Cfg<T, L>::NT), 0, st, a);
    (void)pdiv;
  } else if constexpr (DIM == 3 && sizeof(T) == 8 && L == 768) {
    // FP64 at L = 768: one line pair per workgroup with a single six-sequence
    // forward round, whatever pdiv says (one plan per precision and size)
    const long blocks = a.nlines / 2 * nsamples;
    hipLaunchKernelGGL((euler_row6_kernel<T>), dim3((unsigned)blocks), dim3(Row6Cfg::NT), 0, st, a);
    (void)pdiv;
  } else if constexpr (DIM == 3 && (L == 96 || (L == 192 && sizeof(T) == 4))) {
    constexpr int P0 = 1536 / L, P1 = P0 / 2;
    if (pdiv == 2)
      launch_row4_P<T, L, P1>(a, nsamples, st);
    else
      launch_row4_P<T, L, P0>(a, nsamples, st);
  } else if constexpr (DIM == 3) {
    // 3D at L = 384, and FP64 at L = 192: 768/L line pairs per workgroup with a
    // single forward round, whatever pdiv says (one plan per precision and size)
    const long blocks = a.nlines / (2 * Row6pCfg<L>::P) * nsamples;
    hipLaunchKernelGGL((euler_row6p_kernel<T, L>), dim3((unsigned)blocks), dim3(Row6pCfg<L>::NT), 0, st, a);
    (void)pdiv;
  } else if constexpr (L == 1536) {
    // 2D at L = 1536: one line pair per 192-thread workgroup, 24 values per thread
    const long blocks = a.nlines / 2 * nsamples;
    hipLaunchKernelGGL((euler_row2w_kernel<T>), dim3((unsigned)blocks), dim3(Row2wCfg::NT), 0, st, a);
    (void)pdiv;
  } else if constexpr (L == 768 || L == 384) {
    // 2D at L = 768, 384: 1536/L line pairs per 192-thread workgroup, the 3D forward
    const long blocks = a.nlines / (2 * Row2pCfg<L>::P) * nsamples;
    hipLaunchKernelGGL((euler_row2p_kernel<T, L>), dim3((unsigned)blocks), dim3(Row2pCfg<L>::NT), 0, st, a);
    (void)pdiv;
  } else {
    // one shape: 768/L line pairs, 192 threads
    const long blocks = a.nlines / (2 * Row2Cfg<T, L>::P) * nsamples;
    hipLaunchKernelGGL((euler_row2_kernel<T, L>), dim3((unsigned)blocks), dim3(Row2Cfg<T, L>::NT), 0, st, a);
    (void)pdiv;
  }
}

template <typename T, int DIM>
int launch_row_T(int L, const RowArgs& a, long nsamples, hipStream_t st, int pdiv) {
  switch (L) {
    case 96: launch_row_L<T, 96, DIM>(a, nsamples, st, pdiv); return 0;
    case 192:
      if (DIM == 3 && sizeof(T) == 8 && a.nlines % 8 != 0) return 2;  // four line pairs per workgroup
      launch_row_L<T, 192, DIM>(a, nsamples, st, pdiv);
      return 0;
    case 384:
      if (a.nlines % (DIM == 3 ? 4 : 8) != 0) return 2;  // two (3D) or four (2D) line pairs per workgroup
      launch_row_L<T, 384, DIM>(a, nsamples, st, pdiv);
      return 0;
    case 768:
      if constexpr ((DIM == 3 && sizeof(T) == 4) || DIM == 2) {
        if (a.nlines % 4 != 0) return 2;  // two line pairs per workgroup
      }
      if (a.nlines % 2 != 0) return 2;
      launch_row_L<T, 768, DIM>(a, nsamples, st, pdiv);
      return 0;
    case 1536:
      if constexpr (DIM == 2) {
        launch_row_L<T, 1536, 2>(a, nsamples, st, pdiv);
        return 0;
      }
      return 1;
    default: return 1;
  }
}

}  // namespace

int launch_row_fused(int tbytes, int dim, int L, const RowArgs& a, long nsamples, hipStream_t st, int pdiv) {
  if (tbytes == 8) return dim == 3 ? launch_row_T<double, 3>(L, a, nsamples, st, pdiv)
                                   : launch_row_T<double, 2>(L, a, nsamples, st, pdiv);
  return dim == 3 ? launch_row_T<float, 3>(L, a, nsamples, st, pdiv)
                  : launch_row_T<float, 2>(L, a, nsamples, st, pdiv);
}

void launch_kspace(int tbytes, int dim, const KArgs& k, unsigned blocks, unsigned samples, hipStream_t st) {
  const dim3 grid(blocks, samples), block(256);
  if (tbytes == 8) {
    if (dim == 3)
      hipLaunchKernelGGL((kspace_kernel<double, 3>), grid, block, 0, st, k);
    else
      hipLaunchKernelGGL((kspace_kernel<double, 2>), grid, block, 0, st, k);
  } else {
    if (dim == 3)
      hipLaunchKernelGGL((kspace_kernel<float, 3>), grid, block, 0, st, k);
    else
      hipLaunchKernelGGL((kspace_kernel<float, 2>), grid, block, 0, st, k);
  }
}

}  // namespace azb
