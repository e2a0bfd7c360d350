// Launchers of the fused row kernels and the k-space kernel.  They live in their
// own translation unit (azb_fused.hip) so that it can be compiled with the
// instruction scheduler that suits these kernels (azeban_amd/build.py).
#pragma once

#include <hip/hip_runtime.h>

#include "azb_kernels.hpp"

namespace azb {

// Fused row kernel (inverse, max|u|, products, forward) over the lines of a:
// tbytes = sizeof(working real), dim 2 or 3, L the padded line length, pdiv the
// line-pair divisor of the 3D kernel.  Returns 0; 1 when no kernel is built for
// (dim, L); 2 when the line count is not a multiple of the kernel's line-pair
// tile (the caller checks the tile that pdiv selects; the FP32 kernel at
// L = 768 always takes two pairs).
int launch_row_fused(int tbytes, int dim, int L, const RowArgs& a, long nsamples, hipStream_t st, int pdiv);

// k-space kernel with the RK stage update over blocks x samples workgroups of 256.
void launch_kspace(int tbytes, int dim, const KArgs& k, unsigned blocks, unsigned samples, hipStream_t st);

}  // namespace azb
