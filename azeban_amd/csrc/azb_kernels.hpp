// Kernels of the spectral-viscosity Euler right-hand side and the SSP-RK3 stage
// updates.  One RHS evaluation (IncompressibleEuler::dudt) is, on the M = 3N/2
// dealiasing grid:
//
//   3D: K1 col  inverse axis 0, rows N -> M (zero padding implicit)
//       K2 col  inverse axis 1, rows N -> M
//       K3 row  c2r along axis 2, B = n^2 u_i u_j, max|u|, r2c along axis 2,
//               only the k <= N/2 outputs kept (truncation implicit)
//       K4 col  forward axis 1, rows M -> N kept
//       K5 col  forward axis 0, rows M -> N kept  => B_hat on the N-grid
//       K6 k-space: i k.B_hat, Leray projection, SmoothCutoff viscosity, fused with
//               the RK stage update
//   2D: K1 col axis 0, K2 row axis 1, K3 col axis 0, K4 k-space.
//
// Reference behaviour restated here: copy_to_padded (copy_to_padded_functions.hpp:
// 11-63), the FFTW/cuFFT c2r convention (complex passes on the leading axes, then
// c2r on the last axis dropping Im at k_last = 0), computeB (incompressible_euler.hpp:
// 168-228, incompressible_euler_functions.hpp:26-45), max_norm (norm.hpp:72-91),
// computeDudt (incompressible_euler.hpp:231-347, incompressible_euler_functions.hpp:
// 47-112), SmoothCutoff1D (spectral_viscosity.hpp:56-77) and SSP_RK3 (ssp_rk3.hpp:48-64).
#pragma once
#include "azb_fft.hpp"

namespace azb {

constexpr int kBins = 64;

// Device-resident scalars of one solver.  The time loop keeps t, dt and the CFL
// maximum on the device so a whole step (or many) runs without a host round trip.
struct DevState {
  unsigned long long umax_bins[4][kBins];  // per-RHS-slot partial max |u| (as double bits)
  double u_max;        // max|u| / M^d of the latest RHS evaluation
  double dt;           // dt of the current step
  double max_dt;       // clamp of the current step
  double time;         // simulation time (device-driven loop)
  double t_end;        // target time of the device-driven loop
  double max_dt_visc;  // C * 2 / (eps (pi N)^2)
  int done;            // set when time >= t_end: all step kernels return at once
  int steps;
  int pad[2];
};

enum ColMode { COL_PAD_IN = 0, COL_TRUNC_OUT = 1, COL_FULL = 2 };
enum KMode { K_DUDT = 0, K_RK1 = 1, K_RK2 = 2, K_RK3 = 3 };

// E (complex values per thread) and radix plan per transform length
template <int L>
struct LenPlan;
template <> struct LenPlan<48> { static constexpr int E = 24; using P = Plan<4, 12>; };
template <> struct LenPlan<96> { static constexpr int E = 24; using P = Plan<8, 12>; };
template <> struct LenPlan<192> { static constexpr int E = 24; using P = Plan<8, 24>; };
template <> struct LenPlan<384> { static constexpr int E = 24; using P = Plan<8, 4, 12>; };
template <> struct LenPlan<768> { static constexpr int E = 24; using P = Plan<8, 8, 12>; };
template <> struct LenPlan<1536> { static constexpr int E = 24; using P = Plan<8, 8, 24>; };
template <> struct LenPlan<16> { static constexpr int E = 16; using P = Plan<16>; };
template <> struct LenPlan<32> { static constexpr int E = 16; using P = Plan<8, 4>; };
template <> struct LenPlan<64> { static constexpr int E = 16; using P = Plan<8, 8>; };
template <> struct LenPlan<128> { static constexpr int E = 16; using P = Plan<8, 16>; };
template <> struct LenPlan<256> { static constexpr int E = 16; using P = Plan<16, 16>; };
template <> struct LenPlan<512> { static constexpr int E = 16; using P = Plan<8, 8, 8>; };
template <> struct LenPlan<1024> { static constexpr int E = 16; using P = Plan<16, 8, 8>; };

// Radix plans of the column passes: V = 0 is LenPlan's, the others are the
// alternatives the autotuner may pick (same E, so the same thread layout and LDS
// exchange; different butterflies, twiddles and rounding).
template <int L, int V>
struct ColPlan { using P = typename LenPlan<L>::P; };
template <> struct ColPlan<384, 1> { using P = Plan<12, 4, 8>; };
template <> struct ColPlan<384, 2> { using P = Plan<4, 4, 24>; };
template <> struct ColPlan<768, 1> { using P = Plan<12, 8, 8>; };
template <> struct ColPlan<768, 2> { using P = Plan<4, 8, 24>; };
template <> struct ColPlan<1536, 1> { using P = Plan<24, 8, 8>; };
template <> struct ColPlan<1536, 2> { using P = Plan<8, 24, 8>; };
constexpr int col_plan_count(int L) { return (L == 384 || L == 768 || L == 1536) ? 3 : 1; }

template <int... Rs>
constexpr int first_radix(Plan<Rs...>) { return PlanFirst<Rs...>::value; }
template <int... Rs>
constexpr int last_radix(Plan<Rs...>) { return plan_last<Rs...>(); }

// ---------------------------------------------------------------------------
// Column FFT: batched complex FFT of length L along a strided axis.  The
// workgroup owns TS consecutive columns (contiguous in memory) of one batch and
// all L rows.  PAD_IN reads only the N nonzero rows of the zero-padded input;
// TRUNC_OUT writes only the N rows the N-grid keeps.
// ---------------------------------------------------------------------------
// Columns are addressed as (batch, q, k): k in [0, nk) is contiguous in memory,
// q in [0, nq) steps by in_qs / out_qs.  A workgroup owns TS consecutive k of one
// (batch, q), so with a q stride that is a multiple of the tile width every
// segment it moves is cache-line aligned.
struct ColArgs {
  const void* in;
  void* out;
  long nq;        // q rows per batch
  int nk;         // valid columns per q row
  int ktiles;     // ceil(nk / TS)
  int nsmall;     // N (unpadded length of the pruned side)
  long in_qs, out_qs;   // q strides (elements)
  long in_rs, out_rs;   // row strides along the transformed axis (elements)
  long in_bs, out_bs;   // batch strides (elements)
  int bps;              // batches per sample (samples are stacked batch-major)
  const void* tw;       // twiddles for L
  const DevState* st;   // one per sample
};

// columns per workgroup: 128-byte segments at working precision
template <typename T, int L>
constexpr int col_ts() {
  return (L % 3 == 0) ? ((6144 * (8 / (int)sizeof(T))) / L < 1 ? 1 : (6144 * (8 / (int)sizeof(T))) / L)
                      : ((4096 * (8 / (int)sizeof(T))) / L < 1 ? 1 : (4096 * (8 / (int)sizeof(T))) / L);
}

// Workgroups are handed to the 8 XCDs round-robin (workgroup b to XCD b mod 8),
// and each XCD has its own L2.  Renumbering so that each XCD takes one contiguous
// range of logical workgroups puts neighbouring column tiles (which share the
// 128-byte lines that straddle their boundary when rows are not line-aligned, as
// in the caller's N/2+1-wide state) in the same L2.  A bijection for any count.
__device__ __forceinline__ unsigned xcd_contiguous(unsigned b, unsigned nb) {
  constexpr unsigned X = 8;
  const unsigned x = b % X, i = b / X, q = nb / X, r = nb % X;
  return x * q + (x < r ? x : r) + i;
}

template <typename Tin, typename Tout, typename T, int L, int TS, int DIR, int MODE, int V = 0>
__global__ __launch_bounds__(TS * L / LenPlan<L>::E, (sizeof(T) == 4 ? 4 : (L == 768 || L == 384 ? 3 : 2))) void col_fft_kernel(ColArgs a) {
  using LP = LenPlan<L>;
  using PP = typename ColPlan<L, V>::P;
  constexpr int E = LP::E;
  constexpr int NT = TS * L / E;
  // exchange in one barrier pair unless the interleaved buffer would exceed 64 KB
  constexpr bool SPLIT = TS * L * 2 * (int)sizeof(T) > 65536;
  // FP32 at L = 768 in the RHS: twiddles from the plan's per-pass table (the
  // launcher points a.tw at it), one base address per butterfly and 16-byte loads
  constexpr bool PT = sizeof(T) == 4 && L == 768 && MODE != COL_FULL;
  using F = BlockFFT2<T, L, TS, NT, DIR, SPLIT, 1, MAP_STD, MAP_STD, false, 1, 0u, 0, 0u, void, PT>;
  constexpr int R0 = first_radix(PP{});
  constexpr int RL = last_radix(PP{});
  __shared__ T lds[ExchangeSize<T, L, TS, SPLIT>::words];
  const int t = threadIdx.x;
  const unsigned blk = xcd_contiguous(blockIdx.x, gridDim.x);
  const int kt = blk % a.ktiles;
  const long rest = blk / a.ktiles;
  const long q = rest % a.nq;
  const long batch = rest / a.nq;
  if (a.st[batch / a.bps].done) return;
  const long c0 = (long)kt * TS;
  // element offsets from the tile's base stay below 2^32 (checked at launch)
  const void* in = (const cx<Tin>*)a.in + batch * a.in_bs + q * a.in_qs + c0;
  void* out = (cx<Tout>*)a.out + batch * a.out_bs + q * a.out_qs + c0;
  const unsigned irs = (unsigned)a.in_rs, ors = (unsigned)a.out_rs;
  const unsigned ncols_left = (unsigned)(a.nk - c0);
  // pruned length of the padded/truncated side (M = 3N/2, checked at launch)
  constexpr unsigned N = MODE == COL_FULL ? L : 2 * L / 3, half = N / 2;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;

  cx<T> v[E];
  // all first-pass loads are issued before any use (maximum bytes in flight)
  F::template load<R0>(v, t, [&](unsigned s, unsigned pos) -> cx<T> {
    unsigned row = pos;
    bool valid = s < ncols_left;
    if (MODE == COL_PAD_IN) {
      if (pos > half) {
        if (pos >= L - half + 1)
          row = pos - (L - N);
        else
          valid = false;
      }
    }
    cx<T> r = mk(T(0), T(0));
    if (valid) {
      const cx<Tin> x = gload<Tin>(in, row * irs + s);
      r = mk(T(x.x), T(x.y));
    }
    return r;
  });
  F::run(v, t, lds, tw, PP{});
  F::template store<RL>(v, t, [&](unsigned s, unsigned pos, cx<T> val) {
    unsigned row = pos;
    bool valid = s < ncols_left;
    if (MODE == COL_TRUNC_OUT) {
      if (pos > half) {
        if (pos >= L - half + 1)
          row = pos - (L - N);
        else
          valid = false;
      }
    }
    if (valid) gstore<Tout>(out, row * ors + s, Tout(val.x), Tout(val.y));
  });
}

// ---------------------------------------------------------------------------
// Row kernel: the last axis of the RHS in one pass.  Each workgroup owns P pairs
// of lines (a, b).  Inverse: the N/2+1 stored modes of each component are
// Hermitian-extended onto the M-point line (Im of the k=0 entry dropped, as the
// reference's c2r does) and two real lines are packed per complex FFT.  The
// products B_m are formed in registers, packed two lines per forward FFT, and
// separated into the two half spectra; only k <= N/2 is written.
// ---------------------------------------------------------------------------
struct RowArgs {
  const void* in;   // [c][line][ls]: N/2+1 modes per line
  void* out;        // [m][line][ls]
  long nlines;      // lines per component (M for 2D, M*M for 3D)
  long ls;          // line stride (elements, >= N/2+1)
  long in_ss, out_ss;  // sample strides (elements)
  double norm;      // 1 / (N^{d/2} M^{d/2})
  const void* tw;
  DevState* st;
  int slot;         // umax slot
};

template <typename T>
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}


// 16-byte load that bypasses the caches' retention.  The k-space kernel reads
// each input once and writes its output with non-temporal stores: measured 2-3 %
// faster (FP32-mixed: 7 %) than cached accesses; in the row kernel non-temporal
// loads measured slower.
template <typename T>
__device__ __forceinline__ cx<T> ntload(const void* p) {
  const vec2_t<T> v = __builtin_nontemporal_load((const vec2_t<T>*)p);
  cx<T> r;
  r.x = v.x;
  r.y = v.y;
  return r;
}

// ---------------------------------------------------------------------------
// K-space kernel with the RK stage update fused.
// ---------------------------------------------------------------------------
struct KArgs {
  const void* bh;        // [m][N-grid] at working precision
  const double2* u_in;   // stage input u_hat
  const double2* u0;     // RK base state (stages 2, 3)
  double2* out;
  long nmodes;           // modes per component
  int n, nh;
  int nhp;               // last-axis stride of B_hat (>= nh)
  long bstride;          // component stride of B_hat
  double eps, s2, k0;    // SmoothCutoff1D with s2 = 2s
  int mode;
  DevState* st;          // one per sample
  // fold = 1 (2D): the max|u| reduction of this RHS runs here instead of in a kernel
  // of its own: u_max = max(bins[slot]) / Md, at stage 1 dt = min(C / (Nd u_max), max_dt)
  int fold, slot;
  double Md, C, Nd;
};

__device__ __forceinline__ double sv_eval(double k, double eps, double s2, double k0) {
  const double k1 = fabs(k) / k0;
  const double k2 = k1 * k1;
  const double k4 = k2 * k2;
  const double k8 = k4 * k4;
  const double k16 = k8 * k8;
  const double k18 = k16 * k2;
  const double qk = 1 - exp(-k18);
  const double kp = s2 == 2.0 ? k * k : pow(k, s2);
  return -eps * kp * qk;
}

__device__ __forceinline__ double2 cmuli(double k, double2 b) {
  // complex_t(0, k) * b
  return make_double2(0.0 * b.x - k * b.y, 0.0 * b.y + k * b.x);
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void kspace_kernel(KArgs a) {
  // grid.y = sample; samples are stacked with strides DIM*nmodes (state) and
  // 3(DIM-1)*bstride (B_hat)
  const long sample = blockIdx.y;
  DevState* st = a.st + sample;
  if (st->done) return;
  // stage step: from the device state, or (fold) from the row kernel's max|u| bins,
  // which every wave reduces itself at stage 1 (the first workgroup of a sample in
  // every mode, to store u_max); the same operations as finalize_umax
  double dt;
  if (a.fold) {
    const bool first = blockIdx.x == 0;
    double um = 0.0;
    if (a.mode == K_RK1 || first) {
      const double v = __longlong_as_double((long long)st->umax_bins[a.slot][threadIdx.x & (kBins - 1)]);
      um = wave_max<double>(v) / a.Md;
    }
    dt = a.mode == K_RK1 ? fmin(a.C / (a.Nd * um), st->max_dt) : st->dt;
    if (first && threadIdx.x == 0) {
      st->u_max = um;
      if (a.mode == K_RK1) st->dt = dt;
    }
  } else {
    dt = st->dt;
  }
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.nmodes) return;
  const double twopi = 2.0 * 3.14159265358979323846;
  const int N = a.n, NH = a.nh;
  const cx<T>* bh = (const cx<T>*)a.bh + sample * (3 * (DIM - 1)) * a.bstride;
  // 32-bit index arithmetic: the supported grids have fewer than 2^27 modes per component
  const unsigned ui = (unsigned)idx;
  const int kk = (int)(ui % (unsigned)NH);
  const unsigned rest = ui / (unsigned)NH;
  const long bidx = (long)rest * a.nhp + kk;
  auto B = [&](int m) -> double2 {
    const cx<T> z = ntload<T>(bh + (long)m * a.bstride + bidx);
    return make_double2((double)z.x, (double)z.y);
  };
  double absk2;
  double2 L[3];
  if (DIM == 2) {
    const int i = (int)rest;
    const int i_ = i >= N / 2 + 1 ? i - N : i;
    const double k1 = twopi * i_;
    const double k2 = twopi * kk;
    absk2 = k1 * k1 + k2 * k2;
    const double2 B11 = B(0), B12 = B(1), B22 = B(2);
    const double2 t1 = cmuli(k1, B11), t2 = cmuli(k2, B12);
    const double2 b1 = make_double2(t1.x + t2.x, t1.y + t2.y);
    const double2 t3 = cmuli(k1, B12), t4 = cmuli(k2, B22);
    const double2 b2 = make_double2(t3.x + t4.x, t3.y + t4.y);
    const double inv = 1.0 / absk2;
    const double k11 = k1 * k1 * inv, k12 = k1 * k2 * inv, k22 = k2 * k2 * inv;
    const double c11 = 1.0 - k11, c12 = 0.0 - k12, c22 = 1.0 - k22;
    L[0] = make_double2(c11 * b1.x + c12 * b2.x, c11 * b1.y + c12 * b2.y);
    L[1] = make_double2(c12 * b1.x + c22 * b2.x, c12 * b1.y + c22 * b2.y);
  } else {
    const int j = (int)(rest % (unsigned)N);
    const int i = (int)(rest / (unsigned)N);
    const int i_ = i >= N / 2 + 1 ? i - N : i;
    const int j_ = j >= N / 2 + 1 ? j - N : j;
    const double k1 = twopi * i_, k2 = twopi * j_, k3 = twopi * kk;
    absk2 = k1 * k1 + k2 * k2 + k3 * k3;
    const double2 B11 = B(0), B21 = B(1), B22 = B(2), B31 = B(3), B32 = B(4), B33 = B(5);
    double2 b[3];
    {
      const double2 x = cmuli(k1, B11), y = cmuli(k2, B21), z = cmuli(k3, B31);
      b[0] = make_double2(x.x + y.x + z.x, x.y + y.y + z.y);
    }
    {
      const double2 x = cmuli(k1, B21), y = cmuli(k2, B22), z = cmuli(k3, B32);
      b[1] = make_double2(x.x + y.x + z.x, x.y + y.y + z.y);
    }
    {
      const double2 x = cmuli(k1, B31), y = cmuli(k2, B32), z = cmuli(k3, B33);
      b[2] = make_double2(x.x + y.x + z.x, x.y + y.y + z.y);
    }
    const double inv = 1.0 / absk2;
    const double k11 = k1 * k1 * inv, k12 = k1 * k2 * inv, k13 = k1 * k3 * inv;
    const double k22 = k2 * k2 * inv, k23 = k2 * k3 * inv, k33 = k3 * k3 * inv;
    const double P[3][3] = {{1.0 - k11, 0.0 - k12, 0.0 - k13},
                            {0.0 - k12, 1.0 - k22, 0.0 - k23},
                            {0.0 - k13, 0.0 - k23, 1.0 - k33}};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      L[r] = make_double2(P[r][0] * b[0].x + P[r][1] * b[1].x + P[r][2] * b[2].x,
                          P[r][0] * b[0].y + P[r][1] * b[1].y + P[r][2] * b[2].y);
    }
  }
  const double nu = sv_eval(sqrt(absk2), a.eps, a.s2, a.k0);
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    const long o = (sample * DIM + c) * a.nmodes + idx;
    const cx<double> uu = ntload<double>(a.u_in + o);
    const double2 u = make_double2(uu.x, uu.y);
    double2 d = absk2 == 0 ? make_double2(0.0, 0.0)
                           : make_double2(-L[c].x + nu * u.x, -L[c].y + nu * u.y);
    double2 r;
    if (a.mode == K_DUDT) {
      r = d;
    } else if (a.mode == K_RK1) {
      r = make_double2(1.0 * u.x + dt * d.x, 1.0 * u.y + dt * d.y);
    } else {
      const double2 tmp = make_double2(1.0 * u.x + dt * d.x, 1.0 * u.y + dt * d.y);
      const cx<double> ww = ntload<double>(a.u0 + o);
      const double2 w = make_double2(ww.x, ww.y);
      if (a.mode == K_RK2) {
        r = make_double2(0.75 * w.x + 0.25 * tmp.x, 0.75 * w.y + 0.25 * tmp.y);
      } else {
        r = make_double2((2.0 / 3) * tmp.x + (1.0 / 3) * w.x, (2.0 / 3) * tmp.y + (1.0 / 3) * w.y);
      }
    }
    vec2_t<double> rv;
    rv.x = r.x;
    rv.y = r.y;
    __builtin_nontemporal_store(rv, (vec2_t<double>*)(a.out + o));
  }
}

}  // namespace azb

namespace azb {

// ---------------------------------------------------------------------------
// Plain last-axis real transforms (standalone FFT entry points): pairs of real
// lines share one complex FFT.  c2r drops Im at k = 0 and k = L/2 (FFTW/cuFFT
// convention); r2c writes the L/2+1 non-negative frequencies.
// ---------------------------------------------------------------------------
struct PlainRowArgs {
  const void* in;
  void* out;
  long nlines;
  const void* tw;
};

template <typename T, int L, int P, int DIR>
__global__ __launch_bounds__(2 * P * L / LenPlan<L>::E / 2) void plain_row_kernel(PlainRowArgs a) {
  // TS = P (one complex FFT per pair of lines)
  using LP = LenPlan<L>;
  constexpr int TS = P;
  constexpr int NT = TS * L / LP::E;
  constexpr int NH = L / 2 + 1;
  constexpr bool SPLIT = sizeof(T) == 8;
  constexpr int R0 = first_radix(typename LP::P{});
  constexpr int RL = last_radix(typename LP::P{});
  using F = BlockFFT<T, L, TS, NT, DIR, SPLIT>;
  __shared__ T X[SPLIT ? TS * L : 2 * TS * L];
  __shared__ cx<T> stg[DIR > 0 ? 2 * P * NH : 1];
  const int t = threadIdx.x;
  const long line0 = (long)blockIdx.x * 2 * P;
  const cx<T>* tw = (const cx<T>*)a.tw;
  cx<T> v[LP::E];
  if (DIR > 0) {
    const cx<T>* in = (const cx<T>*)a.in;
    for (int e = t; e < 2 * P * NH; e += NT) {
      const long line = line0 + e / NH;
      stg[e] = line < a.nlines ? in[line * NH + e % NH] : mk(T(0), T(0));
    }
    __syncthreads();
    F::template load<R0>(v, t, [&](int s, int pos) -> cx<T> {
      const cx<T>* xs = stg + (2 * s) * NH;
      const cx<T>* ys = stg + (2 * s + 1) * NH;
      if (pos == 0 || pos == L / 2) return mk(xs[pos].x, ys[pos].x);
      if (pos < L / 2) {
        const cx<T> X0 = xs[pos], Y0 = ys[pos];
        return mk(X0.x - Y0.y, X0.y + Y0.x);
      }
      const cx<T> X0 = xs[L - pos], Y0 = ys[L - pos];
      return mk(X0.x + Y0.y, Y0.x - X0.y);
    });
    F::run(v, t, X, tw, true, typename LP::P{});
    T* out = (T*)a.out;
    F::template store<RL>(v, t, [&](int s, int pos, cx<T> val) {
      const long la = line0 + 2 * s, lb = la + 1;
      if (la < a.nlines) out[la * L + pos] = val.x;
      if (lb < a.nlines) out[lb * L + pos] = val.y;
    });
  } else {
    const T* in = (const T*)a.in;
    F::template load<R0>(v, t, [&](int s, int pos) -> cx<T> {
      const long la = line0 + 2 * s, lb = la + 1;
      const T xa = la < a.nlines ? in[la * L + pos] : T(0);
      const T xb = lb < a.nlines ? in[lb * L + pos] : T(0);
      return mk(xa, xb);
    });
    F::run(v, t, X, tw, true, typename LP::P{});
    cx<T>* out = (cx<T>*)a.out;
    constexpr int items = TS * NH;
    for (int ph = 0; ph < 2; ++ph) {
      __syncthreads();
      F::template store<RL>(v, t, [&](int s, int pos, cx<T> val) {
        X[s * L + pos] = ph == 0 ? val.x : val.y;
      });
      __syncthreads();
      for (int it = t; it < items; it += NT) {
        const int k = it % NH, s = it / NH;
        const T zk = X[s * L + k];
        const T zm = X[s * L + (k == 0 ? 0 : L - k)];
        const long la = line0 + 2 * s, lb = la + 1;
        // stage the first half-result in the output (re-read in phase 1)
        if (ph == 0) {
          if (la < a.nlines) out[la * NH + k].x = T(0.5) * (zk + zm);
          if (lb < a.nlines) out[lb * NH + k].y = T(-0.5) * (zk - zm);
        } else {
          if (la < a.nlines) out[la * NH + k].y = T(0.5) * (zk - zm);
          if (lb < a.nlines) out[lb * NH + k].x = T(0.5) * (zk + zm);
        }
      }
    }
  }
}

}  // namespace azb

namespace azb {

// ---------------------------------------------------------------------------
// 3D row kernel, 12 values per thread.  Per workgroup: P line pairs (a, b),
// TS = 3P packed sequences, NT = P*L/4 threads.
//   inverse  : packed c2r of (u1a + i u2a, u3a + i u3b, u1b + i u2b), run once.  Its
//              first pass (radix 12, L/12 butterflies per sequence, consecutive
//              butterflies in consecutive lanes) loads straight from HBM with
//              contiguous lanes; its last pass (radix 4, MAP_SEQ) leaves u1, u2, u3 of
//              both lines at the same 4 positions in every thread, where they stay
//              in registers for both product rounds;
//   products : round 0 -> (B11, B21, B22), round 1 -> (B31, B32, B33), packed
//              (line a) + i (line b), feeding the forward transform's MAP_SEQ first
//              pass at exactly those positions;
//   forward  : last pass radix 12 with butterfly j of a sequence in lane j of its
//              wave, so the butterfly holding the mirrored positions L - pos
//              (butterfly L/12 - j) is another lane of the same wave: the four mirror
//              values a thread needs for separation come over lane permutes, and
//              the k <= N/2 outputs leave as contiguous 16-byte stores.
// ---------------------------------------------------------------------------
// Packed first inverse pass of a row kernel, from HBM: radix R (12 or 24) with one
// butterfly per thread, butterfly j (< NB = L/R) of a sequence X + iY at positions
// j + NB r, X and Y the half spectra (N/2 + 1 modes) of two real lines.  Positions
// up to N/2 = (R/3) NB read k = pos (r < R/3; r = R/3 only for j = 0), positions
// from L - N/2 = (2R/3) NB read the conjugates at k = L - pos (r >= 2R/3), the rest
// are the zero padding; the k = 0 entry drops Im as the reference's c2r does.  Ends
// with the radix-R DFT.
template <typename T, int NB, int R = 12>
__device__ __forceinline__ void row_first_pass(cx<T>* v, const cx<T>* __restrict__ px,
                                               const cx<T>* __restrict__ py, unsigned j) {
  static_assert(R % 3 == 0, "N/2 = L/3 must fall on a butterfly row");
  constexpr int R3 = R / 3;
  const bool j0 = j == 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (r < R3) {
      const cx<T> X0 = px[j + NB * r], Y0 = py[j + NB * r];
      v[r] = mk(X0.x - Y0.y, X0.y + Y0.x);
      if (r == 0 && j0) v[r] = mk(X0.x, Y0.x);
    } else if (r >= 2 * R3) {
      const unsigned kk = NB * (R - r) - j;
      const cx<T> X0 = px[kk], Y0 = py[kk];
      v[r] = mk(X0.x + Y0.y, Y0.x - X0.y);
    } else {
      v[r] = mk(T(0), T(0));
      if (r == R3 && j0) {
        const cx<T> X0 = px[R3 * NB], Y0 = py[R3 * NB];
        v[r] = mk(X0.x - Y0.y, X0.y + Y0.x);
      }
    }
  }
  DFT<R, +1, T>::run(v);
}

// Separation of a packed forward transform Z = A + iB (A, B spectra of two real
// lines) after a last pass of radix R (12 or 24) with butterfly j of a sequence in
// lane j of its wave or half-wave (NB = L/R dividing 64).  The thread holds
// Z[j + NB r], r < R; the mirror of position
// j + NB r is (NB - j) + NB (R - 1 - r), index R - 1 - r of the lane holding
// butterfly NB - j (for j = 0: index R - r of this lane).  Stores A[k], B[k] for the
// k <= N/2 = (R/3) NB this lane owns (r < R/3, and r = R/3 for j = 0) as contiguous
// non-temporal stores.
template <typename T, int NB, int R = 12>
__device__ __forceinline__ void row_separate_store(const cx<T>* v, unsigned lane, unsigned j, vec2_t<T>* oa,
                                                   vec2_t<T>* ob) {
  constexpr int R3 = R / 3;
  // butterfly j of a sequence in lane j of its wave (or half-wave)
  static_assert(64 % NB == 0, "a sequence's butterflies within one wave");
  const int mlane = (int)(lane - j + (NB - j) % NB);
  cx<T> f[R3];
#pragma unroll
  for (int q = 0; q < R3; ++q) {
    f[q].x = __shfl(v[R - 1 - q].x, mlane);
    f[q].y = __shfl(v[R - 1 - q].y, mlane);
  }
  auto emit = [&](unsigned pos, const cx<T>& zk, const cx<T>& zm) {
    vec2_t<T> va, vb;
    va.x = T(0.5) * (zk.x + zm.x);
    va.y = T(0.5) * (zk.y - zm.y);
    vb.x = T(0.5) * (zk.y + zm.y);
    vb.y = T(-0.5) * (zk.x - zm.x);
    __builtin_nontemporal_store(va, oa + pos);
    __builtin_nontemporal_store(vb, ob + pos);
  };
  const bool j0 = j == 0;
#pragma unroll
  for (int r = 0; r < R3; ++r) {
    cx<T> zm = f[r];
    if (r == 0) {
      if (j0) zm = v[0];
    } else {
      if (j0) zm = f[r - 1];
    }
    emit(j + NB * r, v[r], zm);
  }
  if (j0) emit(R3 * NB, v[R3], f[R3 - 1]);
}

template <int L>
struct Row4Plan;
template <> struct Row4Plan<192> { using Inv = Plan<12, 4, 4>; using Fwd = Plan<4, 4, 12>; };
template <> struct Row4Plan<96> { using Inv = Plan<12, 2, 4>; using Fwd = Plan<4, 2, 12>; };

template <typename T, int L, int P>
struct Row4Cfg {
  static constexpr int TS = 3 * P;
  static constexpr int NT = P * L / 4;
  static constexpr int E = 12;
  static constexpr int N = 2 * L / 3;
  static constexpr int NB = L / 12;  // butterflies per sequence in the radix-12 passes
  static constexpr int SOFF = 11;
  static constexpr int XWORDS = 2 * ExLayout<T, L, TS, true, SOFF>::elems;
};

template <typename T, int L, int P>
__global__ __launch_bounds__(P * L / 4) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 4, sizeof(T) == 8 ? 3 : 4))) void euler_row4_kernel(RowArgs a) {
  using C = Row4Cfg<T, L, P>;
  using PL = Row4Plan<L>;
  constexpr int NT = C::NT, E = C::E, N = C::N, TS = C::TS, NB = C::NB;
  static_assert(64 % NB == 0 && NT % 64 == 0, "a wave must hold whole sequences in the radix-12 passes");
  using FI = BlockFFT2<T, L, TS, NT, +1, false, 3, MAP_SEQ1, MAP_SEQ, true, C::SOFF>;
  using FF = BlockFFT2<T, L, TS, NT, -1, false, 3, MAP_SEQ, MAP_SEQ1, true, C::SOFF>;
  static_assert(FI::E == E && FF::E == E && FI::lds_words == C::XWORDS, "exchange size");
  static_assert(first_radix(typename PL::Inv{}) == 12 && last_radix(typename PL::Inv{}) == 4 &&
                    first_radix(typename PL::Fwd{}) == 4 && last_radix(typename PL::Fwd{}) == 12,
                "plans must start/end on the radices the mappings assume");
  __shared__ T X[C::XWORDS];
  __shared__ double red[NT / 64];
  const long bps = a.nlines / (2 * P);  // workgroups per sample
  const long sample = blockIdx.x / bps;
  DevState* st = a.st + sample;
  if (st->done) return;

  const int t = threadIdx.x;
  const long line0 = (blockIdx.x % bps) * (2 * P);
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + sample * a.in_ss;
  cx<T>* __restrict__ out = (cx<T>*)a.out + sample * a.out_ss;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const T nrm = T(a.norm);

  // first inverse pass from HBM; sequence s is type s % 3 of pair s / 3:
  // (u1a + i u2a, u3a + i u3b, u1b + i u2b)
  cx<T> v[E];
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    const unsigned s = tt / NB, j = tt % NB;
    const unsigned pair = s / 3, type = s % 3;
    const long la = line0 + 2 * pair, lb = la + 1;
    const long lx = type == 0 ? la : (type == 1 ? 2 * a.nlines + la : lb);
    const long ly = type == 0 ? a.nlines + la : (type == 1 ? 2 * a.nlines + lb : a.nlines + lb);
    row_first_pass<T, NB>(v, in + lx * a.ls, in + ly * a.ls, j);
  }
  FI::run_rest(v, t, X, tw, typename PL::Inv{});
  // u[g*4 + r]: packed sequence type g of this thread's pair at its 4 positions
  cx<T> u[E];
  T lm = T(0);  // max|u| at working precision (exact), widened once
#pragma unroll
  for (int e = 0; e < E; ++e) {
    u[e] = v[e];
    lm = fmax(lm, fmax(fabs(v[e].x), fabs(v[e].y)));
  }
  double lmax = (double)lm;
  // settle the max here: left to itself the compiler sinks the reduction past the
  // product rounds and keeps its inputs alive (spilled) across them
  asm volatile("" : "+v"(lmax));

#pragma unroll 1
  for (int round = 0; round < 2; ++round) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // p0 = (u1a, u2a), p1 = (u3a, u3b), p2 = (u1b, u2b)
      const cx<T> p0 = u[r], p1 = u[4 + r], p2 = u[8 + r];
      const T a1 = nrm * p0.x, a2 = nrm * p0.y, b1 = nrm * p2.x, b2 = nrm * p2.y;
      if (round == 0) {
        v[r] = mk(a1 * a1, b1 * b1);      // B11
        v[4 + r] = mk(a2 * a1, b2 * b1);  // B21
        v[8 + r] = mk(a2 * a2, b2 * b2);  // B22
      } else {
        const T a3 = nrm * p1.x, b3 = nrm * p1.y;
        v[r] = mk(a3 * a1, b3 * b1);      // B31
        v[4 + r] = mk(a3 * a2, b3 * b2);  // B32
        v[8 + r] = mk(a3 * a3, b3 * b3);  // B33
      }
    }
    FF::run(v, t, X, tw, typename PL::Fwd{});
    {
      unsigned tt = t;
      asm volatile("" : "+v"(tt));
      const unsigned s = tt / NB, j = tt % NB;
      const unsigned pair = s / 3, m = round * 3 + s % 3;
      const long la = line0 + 2 * pair, lb = la + 1;
      row_separate_store<T, NB>(v, tt & 63, j, (vec2_t<T>*)(out + ((long)m * a.nlines + la) * a.ls),
                                (vec2_t<T>*)(out + ((long)m * a.nlines + lb) * a.ls));
    }
  }

  lmax = wave_max<T>(lmax);
  if ((t & 63) == 0) red[t >> 6] = lmax;
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int w = 0; w < NT / 64; ++w) mx = fmax(mx, red[w]);
    atomicMax(&st->umax_bins[a.slot][blockIdx.x & (kBins - 1)],
              (unsigned long long)__double_as_longlong(mx));
  }
}


// ---------------------------------------------------------------------------
// 3D row kernel for FP32 working precision, 24 values per thread.  Per workgroup:
// two line pairs, six packed sequences (the row-4 kernel's three per pair), 192
// threads.  Inverse plan 24.8.4 and forward plan 4.8.24: two exchanges per
// transform instead of three, and the radix-8 and radix-24 passes apply fewer
// twiddles per point than radix-4 ones.  FP32 arithmetic issues at the FP64 rate
// here (packed FP32 is off), so the row-4 kernel is bound by instruction issue in
// FP32 and this one trades registers (FP32 values take half of FP64's) for
// instructions.  The radix-24 first and last passes keep butterfly j of a sequence
// in lane j of its half-wave (NB = L/24 = 32), so the Hermitian extension and the
// half-spectrum separation stay within a wave as in the row-4 kernel.
// ---------------------------------------------------------------------------
template <int L>
struct Row8Plan;
template <> struct Row8Plan<768> { using Inv = Plan<24, 8, 4>; using Fwd = Plan<4, 8, 24>; };

template <typename T, int L>
struct Row8Cfg {
  static constexpr int P = 2;   // line pairs per workgroup
  static constexpr int TS = 6;  // packed sequences
  static constexpr int NT = 192;
  static constexpr int E = 24;
  static constexpr int N = 2 * L / 3;
  static constexpr int NB = L / 24;
  // sequence stride 5 (mod the 256-byte bank row): a model of the exchanges' lane
  // patterns gives an average conflict degree of 1.58 against 3.55 for stride 1
  static constexpr int SOFF = 5;
  static constexpr int XWORDS = 2 * ExLayout<T, L, TS, true, SOFF>::elems;
  // per-pass twiddle tables (inverse, then forward) after the W_L^e table
  static constexpr int TW_INV = L, TW_FWD = L + pass_table_entries(typename Row8Plan<L>::Inv{});
  static constexpr int TW_END = TW_FWD + pass_table_entries(typename Row8Plan<L>::Fwd{});
};

// Per-pass twiddle tables of the FP32 768-point column passes, after the row
// kernel's: for each radix plan V (ColPlan<768, V>) the forward table, then the
// inverse one.
constexpr int col_ptab_size(int V) {
  return V == 0 ? pass_table_entries(ColPlan<768, 0>::P{})
                : (V == 1 ? pass_table_entries(ColPlan<768, 1>::P{}) : pass_table_entries(ColPlan<768, 2>::P{}));
}
constexpr int col_ptab_offset(int V, int DIR) {
  int off = Row8Cfg<float, 768>::TW_END;
  for (int v = 0; v < V; ++v) off += 2 * col_ptab_size(v);
  return off + (DIR > 0 ? col_ptab_size(V) : 0);
}
constexpr int kTw768F32End = col_ptab_offset(2, +1) + col_ptab_size(2);

template <typename T, int L>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(3, 3))) void euler_row8_kernel(RowArgs a) {
  using C = Row8Cfg<T, L>;
  using PL = Row8Plan<L>;
  constexpr int NT = C::NT, E = C::E, TS = C::TS, NB = C::NB, P = C::P;
  static_assert(64 % NB == 0 && TS * L == NT * E, "thread layout");
  // Twiddles from per-pass tables (one base address per butterfly, 16-byte loads:
  // a third fewer integer instructions than the W_L^e table's per-power indices).
  // The radix-8 passes run in half-waves (MAP_HALF), so the exchanges next to the
  // radix-24 passes stay inside a wave (no workgroup barrier: 7 instead of 13).
  // Swizzled layouts that a model of the banks puts conflict-free for those two
  // exchanges cost about 320 more integer instructions per wave here, where the
  // VALU is the bound, so the padded layout stays.
  using FI = BlockFFT2<T, L, TS, NT, +1, false, 6, MAP_SEQ1, MAP_SEQ, true, C::SOFF, 0u, 0, 0b010u, void, true>;
  using FF = BlockFFT2<T, L, TS, NT, -1, false, 6, MAP_SEQ, MAP_SEQ1, true, C::SOFF, 0u, 0, 0b010u, void, true>;
  static_assert(FI::E == E && FF::E == E && FI::lds_words == C::XWORDS, "exchange size");
  static_assert(first_radix(typename PL::Inv{}) == 24 && last_radix(typename PL::Inv{}) == 4 &&
                    first_radix(typename PL::Fwd{}) == 4 && last_radix(typename PL::Fwd{}) == 24,
                "plans must start/end on the radices the mappings assume");
  __shared__ T X[C::XWORDS];
  __shared__ double red[NT / 64];
  const long bps = a.nlines / (2 * P);  // workgroups per sample
  const long sample = blockIdx.x / bps;
  DevState* st = a.st + sample;
  if (st->done) return;

  const int t = threadIdx.x;
  const long line0 = (blockIdx.x % bps) * (2 * P);
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + sample * a.in_ss;
  cx<T>* __restrict__ out = (cx<T>*)a.out + sample * a.out_ss;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const T nrm = T(a.norm);

  // first inverse pass from HBM; sequence s = 3 pair + type:
  // type 0: u1a + i u2a, type 1: u3a + i u3b, type 2: u1b + i u2b
  cx<T> v[E];
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    const unsigned s = tt / NB, j = tt % NB;
    const unsigned pair = s / 3, type = s % 3;
    const long la = line0 + 2 * pair, lb = la + 1;
    const long lx = type == 0 ? la : (type == 1 ? 2 * a.nlines + la : lb);
    const long ly = type == 0 ? a.nlines + la : (type == 1 ? 2 * a.nlines + lb : a.nlines + lb);
    row_first_pass<T, NB, 24>(v, in + lx * a.ls, in + ly * a.ls, j);
  }
  FI::run_rest(v, t, X, tw + C::TW_INV, typename PL::Inv{});
  // u[s*4 + r]: sequence s at this thread's 4 positions
  cx<T> u[E];
  T lm = T(0);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    u[e] = v[e];
    lm = fmax(lm, fmax(fabs(v[e].x), fabs(v[e].y)));
  }
  double lmax = (double)lm;
  asm volatile("" : "+v"(lmax));

  // both product rounds unrolled: a rolled loop selects each round's inputs with
  // about a hundred register moves per round
#pragma unroll
  for (int round = 0; round < 2; ++round) {
#pragma unroll
    for (int pr = 0; pr < P; ++pr) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // p0 = (u1a, u2a), p1 = (u3a, u3b), p2 = (u1b, u2b) of pair pr
        const cx<T> p0 = u[(3 * pr) * 4 + r], p1 = u[(3 * pr + 1) * 4 + r], p2 = u[(3 * pr + 2) * 4 + r];
        const T a1 = nrm * p0.x, a2 = nrm * p0.y, b1 = nrm * p2.x, b2 = nrm * p2.y;
        cx<T>* w = v + (3 * pr) * 4 + r;
        if (round == 0) {
          w[0] = mk(a1 * a1, b1 * b1);  // B11
          w[4] = mk(a2 * a1, b2 * b1);  // B21
          w[8] = mk(a2 * a2, b2 * b2);  // B22
        } else {
          const T a3 = nrm * p1.x, b3 = nrm * p1.y;
          w[0] = mk(a3 * a1, b3 * b1);  // B31
          w[4] = mk(a3 * a2, b3 * b2);  // B32
          w[8] = mk(a3 * a3, b3 * b3);  // B33
        }
      }
    }
    FF::run(v, t, X, tw + C::TW_FWD, typename PL::Fwd{});
    {
      unsigned tt = t;
      asm volatile("" : "+v"(tt));
      const unsigned s = tt / NB, j = tt % NB;
      const unsigned pair = s / 3, m = round * 3 + s % 3;
      const long la = line0 + 2 * pair, lb = la + 1;
      row_separate_store<T, NB, 24>(v, tt & 63, j, (vec2_t<T>*)(out + ((long)m * a.nlines + la) * a.ls),
                                    (vec2_t<T>*)(out + ((long)m * a.nlines + lb) * a.ls));
    }
  }

  lmax = wave_max<T>(lmax);
  if ((t & 63) == 0) red[t >> 6] = lmax;
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int w = 0; w < NT / 64; ++w) mx = fmax(mx, red[w]);
    atomicMax(&st->umax_bins[a.slot][blockIdx.x & (kBins - 1)],
              (unsigned long long)__double_as_longlong(mx));
  }
}

// ---------------------------------------------------------------------------
// 3D FP64 row kernel at L = 768 with one forward round.  The inverse is the row-4
// kernel's (three packed sequences of one line pair, 12 values per thread, plan
// 12.4.4.4).  All six products B_m (packed line a + i line b) are formed at once
// from the inverse output, so the inverse values need not stay in registers for a
// second round, and one forward transform of six sequences (24 values per thread,
// plan 4.8.24, two exchanges) replaces the row-4 kernel's two rounds of three
// (three exchanges each).  The forward exchanges go in two halves (real,
// imaginary: 8-byte slots, pad period 32) so the exchange buffer keeps the size
// of the inverse's.  The radix-24 last pass keeps butterfly j of sequence m in
// lane j of a half-wave, so the separation of each packed spectrum stays within a
// wave as in the other row kernels.
// ---------------------------------------------------------------------------
struct Row6Cfg {
  static constexpr int L = 768;
  static constexpr int NT = 192;
  static constexpr int EI = 12, EF = 24;
  static constexpr int NBI = L / 12, NBF = L / 24;
  static constexpr int SOFFI = 11, SOFFF = 5;
  using LI = ExLayout<double, L, 3, true, SOFFI>;
  using LF = ExLayout<double, L, 6, true, SOFFF, 32>;
  // Exchange layouts, chosen with a model of the LDS banks (64 banks of 4 bytes;
  // 16-byte accesses served 16 lanes at a time, 8-byte ones 32) over every write
  // and read pattern.  Inverse exchange 0 (radix-12 outputs, lane stride 12
  // positions): the padded layout puts lanes 4 apart on the same banks (3-way);
  // the swizzle pos ^ ((pos >> 4) & 3) spreads them, and the reads (position steps
  // of 192) stay constant offsets.  Forward exchange 1 (radix 8 in half-waves ->
  // radix 24): pos ^ (((pos >> 5) & 7) << 2) is conflict-free both ways, where
  // the padded layout is 3-way on the writes.  Measured with counters at 512^3:
  // half the bank-conflict cycles and 30 % less waiting for LDS issue.
  struct InvLay {
    template <int P>
    using lay = std::conditional_t<P == 0, XorLayout<double, L, 3, 4, 3, 0, LI::LP>, LI>;
    static constexpr int elems = LI::elems;
  };
  struct FwdLay {
    template <int P>
    using lay = std::conditional_t<P == 1, XorLayout<double, L, 6, 5, 7, 2, LF::LP>, LF>;
    static constexpr int elems = LF::elems;
  };
  static constexpr int XI = 2 * LI::elems;
  static constexpr int XF = LF::elems;
  static constexpr int XWORDS = XI > XF ? XI : XF;
};

template <typename T>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(3, 3))) void euler_row6_kernel(RowArgs a) {
  static_assert(sizeof(T) == 8, "FP64 kernel");
  using C = Row6Cfg;
  constexpr int L = C::L, NT = C::NT, EI = C::EI, EF = C::EF, NBI = C::NBI, NBF = C::NBF;
  // inverse middle passes with each sequence in its own wave (MAP_WAVE): two of
  // its three exchanges need no workgroup barrier; the forward's radix-8 pass with
  // each sequence in its own half-wave (MAP_HALF), so its exchange into the
  // radix-24 pass (half-wave per sequence too) needs none either
  using FI = BlockFFT2<T, L, 3, NT, +1, false, 3, MAP_SEQ1, MAP_SEQ, true, C::SOFFI, 0b0110u, 0, 0u,
                       typename C::InvLay>;
  using FF = BlockFFT2<T, L, 6, NT, -1, true, 6, MAP_SEQ, MAP_SEQ1, true, C::SOFFF, 0u, 32, 0b010u,
                       typename C::FwdLay>;
  static_assert(FI::E == EI && FF::E == EF, "thread layout");
  static_assert(FI::lds_words <= C::XWORDS && FF::lds_words <= C::XWORDS, "exchange size");
  __shared__ T X[C::XWORDS];
  __shared__ double red[NT / 64];
  const long bps = a.nlines / 2;  // workgroups per sample
  const long sample = blockIdx.x / bps;
  DevState* st = a.st + sample;
  if (st->done) return;

  const int t = threadIdx.x;
  const long la = (blockIdx.x % bps) * 2, lb = la + 1;
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + sample * a.in_ss;
  cx<T>* __restrict__ out = (cx<T>*)a.out + sample * a.out_ss;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const T nrm = T(a.norm);

  // inverse: sequence s = type: 0: u1a + i u2a, 1: u3a + i u3b, 2: u1b + i u2b
  cx<T> u[EI];
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    const unsigned s = tt / NBI, j = tt % NBI;
    const long lx = s == 0 ? la : (s == 1 ? 2 * a.nlines + la : lb);
    const long ly = s == 0 ? a.nlines + la : (s == 1 ? 2 * a.nlines + lb : a.nlines + lb);
    row_first_pass<T, NBI>(u, in + lx * a.ls, in + ly * a.ls, j);
  }
  FI::run_rest(u, t, X, tw, Plan<12, 4, 4, 4>{});
  T lm = T(0);
#pragma unroll
  for (int e = 0; e < EI; ++e) lm = fmax(lm, fmax(fabs(u[e].x), fabs(u[e].y)));
  double lmax = (double)lm;
  asm volatile("" : "+v"(lmax));

  // products at this thread's 4 positions: w[m*4 + r] = B_m (line a) + i B_m (line b),
  // m = 0..5: B11, B21, B22, B31, B32, B33
  cx<T> w[EF];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const cx<T> p0 = u[r], p1 = u[4 + r], p2 = u[8 + r];
    const T a1 = nrm * p0.x, a2 = nrm * p0.y, a3 = nrm * p1.x;
    const T b1 = nrm * p2.x, b2 = nrm * p2.y, b3 = nrm * p1.y;
    w[0 * 4 + r] = mk(a1 * a1, b1 * b1);
    w[1 * 4 + r] = mk(a2 * a1, b2 * b1);
    w[2 * 4 + r] = mk(a2 * a2, b2 * b2);
    w[3 * 4 + r] = mk(a3 * a1, b3 * b1);
    w[4 * 4 + r] = mk(a3 * a2, b3 * b2);
    w[5 * 4 + r] = mk(a3 * a3, b3 * b3);
  }
  FF::run(w, t, X, tw, Plan<4, 8, 24>{});
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    const unsigned m = tt / NBF, j = tt % NBF;
    row_separate_store<T, NBF, 24>(w, tt & 63, j, (vec2_t<T>*)(out + ((long)m * a.nlines + la) * a.ls),
                                   (vec2_t<T>*)(out + ((long)m * a.nlines + lb) * a.ls));
  }

  lmax = wave_max<T>(lmax);
  if ((t & 63) == 0) red[t >> 6] = lmax;
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int q = 0; q < NT / 64; ++q) mx = fmax(mx, red[q]);
    atomicMax(&st->umax_bins[a.slot][blockIdx.x & (kBins - 1)],
              (unsigned long long)__double_as_longlong(mx));
  }
}

// ---------------------------------------------------------------------------
// 2D row kernel (L <= 192), the 3D kernel's layout with one product round.
// Per workgroup: P = 768/L line pairs, 192 threads.  Inverse: 2P packed sequences
// (u1 + i u2 of each line, plan 12 ...): the radix-12 first pass (loaded from HBM)
// runs on the first 2P * L/12 = 128 threads, one butterfly each, the later passes
// on all threads with 8 values each; the last pass (radix 4, MAP_SEQ) leaves u1, u2
// of both lines of a pair at the same four positions in one thread.  Products
// (B11, B12, B22) packed (line a) + i (line b) feed the forward transform (3P
// sequences, 12 values per thread) at those positions; its radix-12 last pass
// separates the two half spectra over lane permutes as in the 3D kernel.
// ---------------------------------------------------------------------------
template <int L>
struct Row2Plan;
template <> struct Row2Plan<192> { using Inv = Plan<12, 4, 4>; using Fwd = Plan<4, 4, 12>; };
template <> struct Row2Plan<96> { using Inv = Plan<12, 2, 4>; using Fwd = Plan<4, 2, 12>; };

template <typename T, int L>
struct Row2Cfg {
  static_assert(L <= 192, "L = 384, 768 and 1536 have kernels of their own (euler_row2p_kernel, euler_row2w_kernel)");
  static constexpr int P = 768 / L;
  static constexpr int NT = P * L / 4;
  static constexpr int TSI = 2 * P, TSF = 3 * P;
  static constexpr int NB = L / 12;
  static constexpr int SOFF = 11;
  static constexpr int XI = 2 * ExLayout<T, L, TSI, true, SOFF>::elems;
  static constexpr int XF = 2 * ExLayout<T, L, TSF, true, SOFF>::elems;
  static constexpr int XWORDS = XI > XF ? XI : XF;
};

template <typename T, int L>
__global__ __launch_bounds__((Row2Cfg<T, L>::NT)) void euler_row2_kernel(RowArgs a) {
  using C = Row2Cfg<T, L>;
  using PL = Row2Plan<L>;
  constexpr int NT = C::NT, P = C::P, NB = C::NB, TSI = C::TSI, TSF = C::TSF;
  using FI = BlockFFT2<T, L, TSI, NT, +1, false, 2, MAP_STD, MAP_SEQ, true, C::SOFF>;
  constexpr int MAPL = MAP_SEQ1;
  using FF = BlockFFT2<T, L, TSF, NT, -1, false, 3, MAP_SEQ, MAPL, true, C::SOFF>;
  using LY = ExLayout<T, L, TSF, true, C::SOFF>;
  static_assert(FI::E == 8 && FF::E == 12, "thread layout");
  static_assert(64 % NB == 0 && TSI * NB % 64 == 0, "whole waves in the radix-12 passes");
  static_assert(first_radix(typename PL::Inv{}) == 12 && last_radix(typename PL::Inv{}) == 4 &&
                    first_radix(typename PL::Fwd{}) == 4 && last_radix(typename PL::Fwd{}) == 12,
                "plans must start/end on the radices the mappings assume");
  __shared__ T X[C::XWORDS];
  __shared__ double red[NT / 64];
  const long bps = a.nlines / (2 * P);  // workgroups per sample
  const long sample = blockIdx.x / bps;
  DevState* st = a.st + sample;
  if (st->done) return;

  const int t = threadIdx.x;
  const long line0 = (blockIdx.x % bps) * (2 * P);
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + sample * a.in_ss;
  cx<T>* __restrict__ out = (cx<T>*)a.out + sample * a.out_ss;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const T nrm = T(a.norm);

  cx<T> v[12];
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    if (tt < (unsigned)(TSI * NB)) {
      const unsigned s = tt / NB, j = tt % NB;  // sequence s: u1 + i u2 of line line0 + s
      const long l = line0 + s;
      row_first_pass<T, NB>(v, in + l * a.ls, in + (a.nlines + l) * a.ls, j);
      // Stockham pass-0 outputs of butterfly j sit at positions 12 j + r
#pragma unroll
      for (int r = 0; r < 12; ++r) {
        const unsigned idx = LY::at(s, 12 * j + r);
        X[2 * idx] = v[r].x;
        X[2 * idx + 1] = v[r].y;
      }
    }
  }
  __syncthreads();
  FI::run_after_first(v, t, X, tw, typename PL::Inv{});

  // v[g*4 + r]: u1 + i u2 of line g of this thread's pair at its 4 positions
  T lm = T(0);
  cx<T> w[12];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const cx<T> pa = v[r], pb = v[4 + r];
    lm = fmax(lm, fmax(fmax(fabs(pa.x), fabs(pa.y)), fmax(fabs(pb.x), fabs(pb.y))));
    const T a1 = nrm * pa.x, a2 = nrm * pa.y, b1 = nrm * pb.x, b2 = nrm * pb.y;
    w[r] = mk(a1 * a1, b1 * b1);      // B11
    w[4 + r] = mk(a1 * a2, b1 * b2);  // B12
    w[8 + r] = mk(a2 * a2, b2 * b2);  // B22
  }
  double lmax = (double)lm;
  FF::run(w, t, X, tw, typename PL::Fwd{});
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    unsigned s, j;
    bmap<MAPL, TSF, NT, L, 12, 3>(tt, 0, s, j);
    const unsigned pair = s / 3, m = s % 3;
    const long la = line0 + 2 * pair, lb = la + 1;
    row_separate_store<T, NB>(w, tt & 63, j, (vec2_t<T>*)(out + ((long)m * a.nlines + la) * a.ls),
                              (vec2_t<T>*)(out + ((long)m * a.nlines + lb) * a.ls));
  }

  lmax = wave_max<T>(lmax);
  if ((t & 63) == 0) red[t >> 6] = lmax;
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int q = 0; q < NT / 64; ++q) mx = fmax(mx, red[q]);
    atomicMax(&st->umax_bins[a.slot][blockIdx.x & (kBins - 1)],
              (unsigned long long)__double_as_longlong(mx));
  }
}

// ---------------------------------------------------------------------------
// 2D row kernel at L = 1536 (N = 1024) with 24 values per thread in the forward.
// One line pair per workgroup of 192 threads.  Inverse (plan 24.8.8): the radix-24
// first pass from HBM on 128 threads (butterfly j of sequence s, NB = 64), written
// to LDS, then two radix-8 passes on all 192 threads with 16 values each; the last
// leaves u1 + i u2 of both lines at the same eight positions t + 192 r in thread t.
// Forward (plan 8.8.24, three sequences B11, B12, B22 packed line a + i line b, 24
// values per thread, exchanges in real/imaginary halves): the radix-8 middle pass
// keeps each sequence in its own wave, so the exchange into the radix-24 pass (one
// sequence per wave, butterfly j in lane j) needs no workgroup barrier, and the
// separation of each packed spectrum stays within a wave.  Two exchanges per
// transform against three (inverse) and four (forward) with 8 and 12 values per
// thread: fewer LDS round trips and barriers in a launch whose 768 workgroups are
// all resident at once, where a workgroup's own chain sets the time.
// ---------------------------------------------------------------------------
struct Row2wCfg {
  static constexpr int L = 1536;
  static constexpr int NT = 192;
  static constexpr int NB = L / 24;  // radix-24 butterflies per sequence
  static constexpr int SOFF = 11;
  using LI = ExLayout<double, L, 2, true, SOFF>;
  using LF = ExLayout<double, L, 3, true, SOFF>;
  using PI = Plan<24, 8, 8>;
  using PF = Plan<8, 8, 24>;
};

template <typename T>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(3))) void euler_row2w_kernel(RowArgs a) {
  using C = Row2wCfg;
  constexpr int L = C::L, NT = C::NT, NB = C::NB;
  // forward exchange halves in FP64 (8-byte slots) keep the buffer at the inverse's
  using FI = BlockFFT2<T, L, 2, NT, +1, false, 2, MAP_STD, MAP_SEQ, true, C::SOFF>;
  using FF = BlockFFT2<T, L, 3, NT, -1, sizeof(T) == 8, 3, MAP_SEQ, MAP_SEQ1, true, C::SOFF, 0b010u>;
  using LY = ExLayout<T, L, 2, true, C::SOFF>;
  static_assert(FI::E == 16 && FF::E == 24, "thread layout");
  constexpr int XWORDS = FI::lds_words > FF::lds_words ? FI::lds_words : FF::lds_words;
  __shared__ T X[XWORDS];
  __shared__ double red[NT / 64];
  const long bps = a.nlines / 2;  // workgroups per sample
  const long sample = blockIdx.x / bps;
  DevState* st = a.st + sample;
  if (st->done) return;

  const int t = threadIdx.x;
  const long la = (blockIdx.x % bps) * 2, lb = la + 1;
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + sample * a.in_ss;
  cx<T>* __restrict__ out = (cx<T>*)a.out + sample * a.out_ss;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const T nrm = T(a.norm);

  cx<T> v[24];
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    if (tt < (unsigned)(2 * NB)) {
      const unsigned s = tt / NB, j = tt % NB;  // sequence s: u1 + i u2 of line la + s
      const long l = la + s;
      row_first_pass<T, NB, 24>(v, in + l * a.ls, in + (a.nlines + l) * a.ls, j);
      // Stockham pass-0 outputs of butterfly j sit at positions 24 j + r
#pragma unroll
      for (int r = 0; r < 24; ++r) {
        const unsigned idx = LY::at(s, 24 * j + r);
        X[2 * idx] = v[r].x;
        X[2 * idx + 1] = v[r].y;
      }
    }
  }
  __syncthreads();
  FI::run_after_first(v, t, X, tw, typename C::PI{});

  // v[g*8 + r]: u1 + i u2 of line g at positions t + 192 r
  T lm = T(0);
  cx<T> w[24];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const cx<T> pa = v[r], pb = v[8 + r];
    lm = fmax(lm, fmax(fmax(fabs(pa.x), fabs(pa.y)), fmax(fabs(pb.x), fabs(pb.y))));
    const T a1 = nrm * pa.x, a2 = nrm * pa.y, b1 = nrm * pb.x, b2 = nrm * pb.y;
    w[r] = mk(a1 * a1, b1 * b1);       // B11
    w[8 + r] = mk(a1 * a2, b1 * b2);   // B12
    w[16 + r] = mk(a2 * a2, b2 * b2);  // B22
  }
  double lmax = (double)lm;
  asm volatile("" : "+v"(lmax));
  FF::run(w, t, X, tw, typename C::PF{});
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    const unsigned m = tt / NB, j = tt % NB;
    row_separate_store<T, NB, 24>(w, tt & 63, j, (vec2_t<T>*)(out + ((long)m * a.nlines + la) * a.ls),
                                  (vec2_t<T>*)(out + ((long)m * a.nlines + lb) * a.ls));
  }

  lmax = wave_max<T>(lmax);
  if ((t & 63) == 0) red[t >> 6] = lmax;
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int q = 0; q < NT / 64; ++q) mx = fmax(mx, red[q]);
    atomicMax(&st->umax_bins[a.slot][blockIdx.x & (kBins - 1)],
              (unsigned long long)__double_as_longlong(mx));
  }
}

// ---------------------------------------------------------------------------
// 2D row kernel at L = 768 and 384 (N = 512, 256) with P = 1536/L line pairs per
// 192-thread workgroup and the 3D kernels' forward: 3P sequences (B11, B12, B22 of
// each pair, packed line a + i line b), 24 values per thread, plans 4.8.24 / 4.4.24
// (at L = 768 the radix-8 pass in half-waves; FP64 as euler_row6_kernel, swizzled
// second exchange; FP32 as euler_row8_kernel, per-pass twiddle tables; every
// exchange in real/imaginary halves).  Inverse (plans 24.8.4 / 24.4.4):
// the radix-24 first pass from HBM on 128 threads (2P sequences u1 + i u2 of lines
// 0..2P-1, butterfly j in lane j of an L/24-lane group), written to LDS, then 16
// values per thread on all 192 threads; the radix-4 last pass leaves four
// sequences (two pairs) at positions t % (L/4) + (L/4) r in thread t, where the
// products are formed.  Two exchanges per transform instead of three.
// ---------------------------------------------------------------------------
template <int L_>
struct Row2pCfg {
  static constexpr int L = L_;
  static constexpr int NT = 192;
  static constexpr int P = 1536 / L;  // line pairs per workgroup
  static constexpr int NB = L / 24;   // radix-24 butterflies per sequence
  static constexpr int SOFFI = 11;
  static constexpr int R1 = L == 768 ? 8 : 4;  // second inverse radix
  using PI = Plan<24, R1, 4>;
  using PF = Plan<4, R1, 24>;
};

template <typename T, int L>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 4))) void euler_row2p_kernel(RowArgs a) {
  using C = Row2pCfg<L>;
  constexpr int NT = C::NT, NB = C::NB, P = C::P, R1 = C::R1;
  constexpr bool F32 = sizeof(T) == 4;
  // every exchange in real/imaginary halves: FP64 keeps four workgroups per CU,
  // FP32 (19 KB) five at four waves per SIMD
  constexpr bool SPI = true;
  // per-pass twiddle tables: FP32 at L = 768 (the 3D FP32 row kernel's, same plans)
  constexpr bool PT = F32 && L == 768;
  using FI = BlockFFT2<T, L, 2 * P, NT, +1, SPI, 4, MAP_STD, MAP_SEQ, true, C::SOFFI, 0u, 0, 0u, void, PT>;
  using FF768 = std::conditional_t<
      F32, BlockFFT2<T, L, 6, NT, -1, true, 6, MAP_SEQ, MAP_SEQ1, true, 5, 0u, 0, 0b010u, void, true>,
      BlockFFT2<T, L, 6, NT, -1, true, 6, MAP_SEQ, MAP_SEQ1, true, Row6Cfg::SOFFF, 0u, 32, 0b010u,
                typename Row6Cfg::FwdLay>>;
  using FF384 = BlockFFT2<T, L, 3 * P, NT, -1, true, 6, MAP_SEQ, MAP_SEQ1, true, 5, 0u, F32 ? 0 : 32>;
  using FF = std::conditional_t<L == 768, FF768, FF384>;
  using LY = ExLayout<T, L, 2 * P, true, C::SOFFI>;
  static_assert(std::is_same_v<LY, typename FI::template XL<0>>, "first-pass layout");
  static_assert(FI::E == 16 && FF::E == 24, "thread layout");
  constexpr int XWORDS = FI::lds_words > FF::lds_words ? FI::lds_words : FF::lds_words;
  static_assert(L == 768 || 4 * (XWORDS * (int)sizeof(T) + 3 * 8) <= 163840, "four workgroups per CU");
  __shared__ T X[XWORDS];
  __shared__ double red[NT / 64];
  const long bps = a.nlines / (2 * P);  // workgroups per sample
  const long sample = blockIdx.x / bps;
  DevState* st = a.st + sample;
  if (st->done) return;

  const int t = threadIdx.x;
  const long line0 = (blockIdx.x % bps) * (2 * P);
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + sample * a.in_ss;
  cx<T>* __restrict__ out = (cx<T>*)a.out + sample * a.out_ss;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const cx<T>* twi = PT ? tw + Row8Cfg<float, 768>::TW_INV : tw;
  const cx<T>* twf = PT ? tw + Row8Cfg<float, 768>::TW_FWD : tw;
  const T nrm = T(a.norm);

  cx<T> v[24];
  if constexpr (!SPI) {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    if (tt < (unsigned)(2 * P * NB)) {
      const unsigned s = tt / NB, j = tt % NB;  // sequence s: u1 + i u2 of line line0 + s
      const long l = line0 + s;
      row_first_pass<T, NB, 24>(v, in + l * a.ls, in + (a.nlines + l) * a.ls, j);
      // Stockham pass-0 outputs of butterfly j sit at positions 24 j + r
#pragma unroll
      for (int r = 0; r < 24; ++r) {
        const unsigned idx = LY::at(s, 24 * j + r);
        X[2 * idx] = v[r].x;
        X[2 * idx + 1] = v[r].y;
      }
    }
    __syncthreads();
    FI::run_after_first(v, t, X, twi, typename C::PI{});
  } else {
    // the first pass's outputs go to LDS in two halves (real, then imaginary parts,
    // one T per slot); pass 1 reads each half with its own mapping, as an exchange
    // of the split transform does, then the transform continues from pass 1
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    const bool writer = tt < (unsigned)(2 * P * NB);
    const unsigned s0 = tt / NB, j0 = tt % NB;
    if (writer) {
      const long l = line0 + s0;
      row_first_pass<T, NB, 24>(v, in + l * a.ls, in + (a.nlines + l) * a.ls, j0);
    }
    cx<T> x[16];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (c == 1) __syncthreads();  // every read of the real half is done
      if (writer) {
#pragma unroll
        for (int r = 0; r < 24; ++r) X[LY::at(s0, 24 * j0 + r)] = c == 0 ? v[r].x : v[r].y;
      }
      __syncthreads();
#pragma unroll
      for (int b = 0; b < 16 / R1; ++b) {
        unsigned s1, j1;
        bmap<MAP_STD, 2 * P, NT, L, R1, 4>(tt, b, s1, j1);
        const unsigned a0 = LY::at(s1, j1);
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          const unsigned idx = LY::linear(L / R1) ? a0 + r * LY::delta(L / R1) : LY::at(s1, j1 + r * (L / R1));
          if (c == 0)
            x[b * R1 + r].x = X[idx];
          else
            x[b * R1 + r].y = X[idx];
        }
      }
    }
    FI::template passes<1, 3, 24, R1, 4>(x, t, X, twi, Plan<R1, 4>{});
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = x[e];
  }

  // v[g*4 + r]: u1 + i u2 of line 4 (t / (L/4)) + g (g = 0..3) at positions
  // t % (L/4) + (L/4) r; products of the thread's pair p (lines 2p, 2p + 1 of the
  // four) into its sequences 3p + m, m = B11, B12, B22
  T lm = T(0);
  cx<T> w[24];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const cx<T> pa = v[(2 * p) * 4 + r], pb = v[(2 * p + 1) * 4 + r];
      lm = fmax(lm, fmax(fmax(fabs(pa.x), fabs(pa.y)), fmax(fabs(pb.x), fabs(pb.y))));
      const T a1 = nrm * pa.x, a2 = nrm * pa.y, b1 = nrm * pb.x, b2 = nrm * pb.y;
      w[(3 * p) * 4 + r] = mk(a1 * a1, b1 * b1);      // B11
      w[(3 * p + 1) * 4 + r] = mk(a1 * a2, b1 * b2);  // B12
      w[(3 * p + 2) * 4 + r] = mk(a2 * a2, b2 * b2);  // B22
    }
  }
  double lmax = (double)lm;
  asm volatile("" : "+v"(lmax));
  FF::run(w, t, X, twf, typename C::PF{});
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    const unsigned s = tt / NB, j = tt % NB;
    const unsigned pair = s / 3, m = s % 3;
    const long la = line0 + 2 * pair, lb = la + 1;
    row_separate_store<T, NB, 24>(w, tt & 63, j, (vec2_t<T>*)(out + ((long)m * a.nlines + la) * a.ls),
                                  (vec2_t<T>*)(out + ((long)m * a.nlines + lb) * a.ls));
  }

  lmax = wave_max<T>(lmax);
  if ((t & 63) == 0) red[t >> 6] = lmax;
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int q = 0; q < NT / 64; ++q) mx = fmax(mx, red[q]);
    atomicMax(&st->umax_bins[a.slot][blockIdx.x & (kBins - 1)],
              (unsigned long long)__double_as_longlong(mx));
  }
}

// ---------------------------------------------------------------------------
// 3D row kernel at L = 384 and 192 (N = 256, 128) with one forward round:
// P = 768/L line pairs per 192-thread workgroup.  The inverse is the row-4
// kernel's (3P packed sequences, 12 values per thread, plans 12.2.4.4 / 12.4.4;
// at L = 384 its two middle passes run in half-waves, each sequence in its own
// half-wave, so two of the three exchanges need no workgroup barrier); all six
// products of a pair are formed at once and one 6P-sequence forward transform
// (24 values per thread, plans 4.4.24 / 4.2.24, exchanges in real/imaginary
// halves) replaces two rounds of 3P.  (At L = 96 the per-sequence pad
// slots of 24 sequences would not leave room for four workgroups per CU; the
// row-4 kernel stays there, and for FP32 at L = 192, where its four waves per
// SIMD measured faster.)  The radix-24 last
// pass keeps butterfly j of a sequence in lane j of an L/24-lane group, so the
// separation of each packed spectrum stays within a wave.
// ---------------------------------------------------------------------------
template <int L>
struct Row6pPlan;
template <> struct Row6pPlan<384> { using Inv = Plan<12, 2, 4, 4>; using Fwd = Plan<4, 4, 24>; };
template <> struct Row6pPlan<192> { using Inv = Plan<12, 4, 4>; using Fwd = Plan<4, 2, 24>; };

template <int L_>
struct Row6pCfg {
  static constexpr int L = L_;
  static constexpr int NT = 192;
  static constexpr int P = 768 / L;  // line pairs per workgroup
  static constexpr int NBI = L / 12, NBF = L / 24;
  // sequence strides chosen so that four workgroups' exchange buffers fit a CU
  static constexpr int SOFFI = 5, SOFFF = 5;
  // middle passes in half-waves where a half-wave holds one sequence (L = 384)
  static constexpr unsigned MIDH = L == 384 ? 0b0110u : 0u;
  // forward pad period (elements): one pad slot per 256 bytes at L = 384; at
  // L = 192 every 16 elements, so that 24 sequences' pad slots fit the budget
  template <typename T>
  static constexpr int fwd_pade() { return L == 384 ? (sizeof(T) == 8 ? 32 : 0) : 16; }
  using PI = typename Row6pPlan<L>::Inv;
  using PF = typename Row6pPlan<L>::Fwd;
};

template <typename T, int L>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 4))) void euler_row6p_kernel(RowArgs a) {
  using C = Row6pCfg<L>;
  constexpr int NT = C::NT, NBI = C::NBI, NBF = C::NBF, P = C::P;
  constexpr bool F64 = sizeof(T) == 8;
  using FI = BlockFFT2<T, L, 3 * P, NT, +1, false, 3, MAP_SEQ1, MAP_SEQ, true, C::SOFFI, 0u, 0, C::MIDH>;
  // forward exchanges in real/imaginary halves: FP64 to fit four workgroups per
  // CU, FP32 (20 KB) to run four waves per SIMD
  using FF = BlockFFT2<T, L, 6 * P, NT, -1, true, 6, MAP_SEQ, MAP_SEQ1, true, C::SOFFF, 0u,
                       C::template fwd_pade<T>()>;
  static_assert(FI::E == 12 && FF::E == 24, "thread layout");
  constexpr int XWORDS = FI::lds_words > FF::lds_words ? FI::lds_words : FF::lds_words;
  static_assert(4 * (XWORDS * (int)sizeof(T) + 3 * 8) <= 163840, "four workgroups per CU");
  __shared__ T X[XWORDS];
  __shared__ double red[NT / 64];
  const long bps = a.nlines / (2 * P);  // workgroups per sample
  const long sample = blockIdx.x / bps;
  DevState* st = a.st + sample;
  if (st->done) return;

  const int t = threadIdx.x;
  const long line0 = (blockIdx.x % bps) * (2 * P);
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + sample * a.in_ss;
  cx<T>* __restrict__ out = (cx<T>*)a.out + sample * a.out_ss;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const T nrm = T(a.norm);

  // first inverse pass from HBM; sequence s is type s % 3 of pair s / 3:
  // (u1a + i u2a, u3a + i u3b, u1b + i u2b)
  cx<T> u[12];
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    const unsigned s = tt / NBI, j = tt % NBI;
    const unsigned pair = s / 3, type = s % 3;
    const long la = line0 + 2 * pair, lb = la + 1;
    const long lx = type == 0 ? la : (type == 1 ? 2 * a.nlines + la : lb);
    const long ly = type == 0 ? a.nlines + la : (type == 1 ? 2 * a.nlines + lb : a.nlines + lb);
    row_first_pass<T, NBI>(u, in + lx * a.ls, in + ly * a.ls, j);
  }
  FI::run_rest(u, t, X, tw, typename C::PI{});
  T lm = T(0);
#pragma unroll
  for (int e = 0; e < 12; ++e) lm = fmax(lm, fmax(fabs(u[e].x), fabs(u[e].y)));
  double lmax = (double)lm;
  asm volatile("" : "+v"(lmax));

  // products of this thread's pair (t / (L/4)) at its 4 positions: w[m*4 + r] =
  // B_m (line a) + i B_m (line b), m = 0..5: B11, B21, B22, B31, B32, B33
  cx<T> w[24];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const cx<T> p0 = u[r], p1 = u[4 + r], p2 = u[8 + r];
    const T a1 = nrm * p0.x, a2 = nrm * p0.y, a3 = nrm * p1.x;
    const T b1 = nrm * p2.x, b2 = nrm * p2.y, b3 = nrm * p1.y;
    w[0 * 4 + r] = mk(a1 * a1, b1 * b1);
    w[1 * 4 + r] = mk(a2 * a1, b2 * b1);
    w[2 * 4 + r] = mk(a2 * a2, b2 * b2);
    w[3 * 4 + r] = mk(a3 * a1, b3 * b1);
    w[4 * 4 + r] = mk(a3 * a2, b3 * b2);
    w[5 * 4 + r] = mk(a3 * a3, b3 * b3);
  }
  FF::run(w, t, X, tw, typename C::PF{});
  {
    unsigned tt = t;
    asm volatile("" : "+v"(tt));
    const unsigned s = tt / NBF, j = tt % NBF;
    const unsigned pair = s / 6, m = s % 6;
    const long la = line0 + 2 * pair, lb = la + 1;
    row_separate_store<T, NBF, 24>(w, tt & 63, j, (vec2_t<T>*)(out + ((long)m * a.nlines + la) * a.ls),
                                   (vec2_t<T>*)(out + ((long)m * a.nlines + lb) * a.ls));
  }

  lmax = wave_max<T>(lmax);
  if ((t & 63) == 0) red[t >> 6] = lmax;
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int q = 0; q < NT / 64; ++q) mx = fmax(mx, red[q]);
    atomicMax(&st->umax_bins[a.slot][blockIdx.x & (kBins - 1)],
              (unsigned long long)__double_as_longlong(mx));
  }
}

}  // namespace azb
