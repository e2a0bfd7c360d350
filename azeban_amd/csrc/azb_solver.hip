// Host side of the C ABI: solver lifetime, workspace layout, kernel launch
// sequences of one RHS evaluation / one SSP-RK3 step / the device-driven time
// loop, and the standalone FFT, pad, axpby and max-norm seams.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/azeban_hip.h"
#include "azb_aux_kernels.hpp"
#include "azb_kernels.hpp"
#include "azb_launch.hpp"

using namespace azb;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(AZB_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// device allocation freed when the scope ends, on every return path (hipFree waits
// for work already enqueued on it)
struct DevFree {
  void* p = nullptr;
  DevFree() = default;
  DevFree(const DevFree&) = delete;
  DevFree& operator=(const DevFree&) = delete;
  ~DevFree() {
    if (p) (void)hipFree(p);
  }
};

bool good_len(int L) {
  switch (L) {
    case 48: case 96: case 192: case 384: case 768: case 1536:
    case 16: case 32: case 64: case 128: case 256: case 512: case 1024:
      return true;
    default:
      return false;
  }
}

// line pairs per workgroup of the fused row kernel
int row_p(int L) {
  int p = 1536 / L;
  return p < 1 ? 1 : p;
}

// ---------------------------------------------------------------------------
// twiddle tables exp(-2 pi i t / L), computed in long double on the host
// ---------------------------------------------------------------------------
template <typename T>
std::vector<cx<T>> host_twiddles(int L) {
  std::vector<cx<T>> h(L);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int t = 0; t < L; ++t) {
    const long double a = twopi * (long double)t / (long double)L;
    h[t].x = (T)cosl(a);
    h[t].y = (T)(-sinl(a));
  }
  return h;
}

// The twiddle buffer of a solver or FFT of length L at precision T: the W_L^e
// table, followed for the FP32 768-point row kernel by its per-pass tables.
template <typename T>
size_t twiddle_entries(int L) {
  if (sizeof(T) == 4 && L == 768) return kTw768F32End;
  return (size_t)L;
}

template <typename T>
std::vector<cx<T>> host_tables(int L) {
  std::vector<cx<T>> h = host_twiddles<T>(L);
  h.resize(twiddle_entries<T>(L));
  if constexpr (sizeof(T) == 4) {
    if (L == 768) {
      using C = Row8Cfg<float, 768>;
      fill_pass_table(h.data() + C::TW_INV, h.data(), L, +1, Row8Plan<768>::Inv{});
      fill_pass_table(h.data() + C::TW_FWD, h.data(), L, -1, Row8Plan<768>::Fwd{});
      fill_pass_table(h.data() + col_ptab_offset(0, -1), h.data(), L, -1, ColPlan<768, 0>::P{});
      fill_pass_table(h.data() + col_ptab_offset(0, +1), h.data(), L, +1, ColPlan<768, 0>::P{});
      fill_pass_table(h.data() + col_ptab_offset(1, -1), h.data(), L, -1, ColPlan<768, 1>::P{});
      fill_pass_table(h.data() + col_ptab_offset(1, +1), h.data(), L, +1, ColPlan<768, 1>::P{});
      fill_pass_table(h.data() + col_ptab_offset(2, -1), h.data(), L, -1, ColPlan<768, 2>::P{});
      fill_pass_table(h.data() + col_ptab_offset(2, +1), h.data(), L, +1, ColPlan<768, 2>::P{});
    }
  }
  return h;
}

template <typename T>
int make_twiddles(int L, void** out) {
  const std::vector<cx<T>> h = host_tables<T>(L);
  DevFree buf;
  HIPCHK(hipMalloc(&buf.p, sizeof(cx<T>) * h.size()));
  HIPCHK(hipMemcpy(buf.p, h.data(), sizeof(cx<T>) * h.size(), hipMemcpyHostToDevice));
  *out = buf.p;
  buf.p = nullptr;
  return AZB_OK;
}

// ---------------------------------------------------------------------------
// launch dispatch over the compile-time transform length
// ---------------------------------------------------------------------------
template <typename Tin, typename Tout, typename T, int L, int TS, int DIR, int MODE, int V = 0>
void launch_col_TS(ColArgs a, long batches, hipStream_t st) {
  constexpr int E = LenPlan<L>::E;
  // the FP32 768-point passes of the RHS read their plan's per-pass table
  if constexpr (sizeof(T) == 4 && L == 768 && MODE != COL_FULL) a.tw = (const cx<T>*)a.tw + col_ptab_offset(V, DIR);
  a.ktiles = (a.nk + TS - 1) / TS;
  const long blocks = batches * a.nq * a.ktiles;
  hipLaunchKernelGGL((col_fft_kernel<Tin, Tout, T, L, TS, DIR, MODE, V>), dim3((unsigned)blocks),
                     dim3(TS * L / E), 0, st, a);
}

// tsdiv in {1, 2, 4} selects tiles of TS0, TS0/2 or TS0/4 columns, TS0 being one
// 128-byte line: narrower tiles give small grids more workgroups
template <typename Tin, typename Tout, typename T, int L, int DIR, int MODE, int V>
void launch_col_LV(ColArgs a, long batches, hipStream_t st, int tsdiv) {
  constexpr int TS0 = azb::col_ts<T, L>();
  constexpr int TS1 = TS0 / 2 >= 1 ? TS0 / 2 : 1;
  constexpr int TS2 = TS0 / 4 >= 1 ? TS0 / 4 : 1;
  if (tsdiv >= 4)
    launch_col_TS<Tin, Tout, T, L, TS2, DIR, MODE, V>(a, batches, st);
  else if (tsdiv >= 2)
    launch_col_TS<Tin, Tout, T, L, TS1, DIR, MODE, V>(a, batches, st);
  else
    launch_col_TS<Tin, Tout, T, L, TS0, DIR, MODE, V>(a, batches, st);
}

// plan in [0, col_plan_count(L)) selects the radix plan (ColPlan)
template <typename Tin, typename Tout, typename T, int L, int DIR, int MODE>
void launch_col_L(ColArgs a, long batches, hipStream_t st, int tsdiv, int plan) {
  if constexpr (col_plan_count(L) > 1) {
    if (plan == 1) return launch_col_LV<Tin, Tout, T, L, DIR, MODE, 1>(a, batches, st, tsdiv);
    if (plan == 2) return launch_col_LV<Tin, Tout, T, L, DIR, MODE, 2>(a, batches, st, tsdiv);
  }
  launch_col_LV<Tin, Tout, T, L, DIR, MODE, 0>(a, batches, st, tsdiv);
}

template <typename Tin, typename Tout, typename T, int DIR, int MODE>
int launch_col(int L, ColArgs a, long batches, hipStream_t st, int tsdiv = 1, int plan = 0) {
  if (plan < 0 || plan >= col_plan_count(L))
    return fail(AZB_ERR_ARG, "radix plan " + std::to_string(plan) + " not built for length " + std::to_string(L));
  // the kernels take the pruned length as 2L/3 and address a tile with 32-bit
  // element offsets from its base
  if (MODE != COL_FULL && 3L * a.nsmall != 2L * L)
    return fail(AZB_ERR_UNSUPPORTED, "pruned column passes need M = 3N/2");
  if ((double)(L - 1) * (double)std::max(a.in_rs, a.out_rs) + (double)L >= 4294967296.0)
    return fail(AZB_ERR_UNSUPPORTED, "column strides exceed 32-bit element offsets");
  switch (L) {
#define AZB_CASE(LL) \
  case LL: launch_col_L<Tin, Tout, T, LL, DIR, MODE>(a, batches, st, tsdiv, plan); break;
    AZB_CASE(96) AZB_CASE(192) AZB_CASE(384) AZB_CASE(768) AZB_CASE(1536)
#undef AZB_CASE
    default:
      return fail(AZB_ERR_UNSUPPORTED, "column FFT length " + std::to_string(L) + " not built");
  }
  return AZB_OK;
}

template <typename T, int DIR>
int launch_col_full(int L, ColArgs a, long batches, hipStream_t st) {
  switch (L) {
#define AZB_CASE(LL) \
  case LL: launch_col_TS<T, T, T, LL, azb::col_ts<T, LL>(), DIR, COL_FULL>(a, batches, st); break;
    AZB_CASE(16) AZB_CASE(32) AZB_CASE(64) AZB_CASE(128) AZB_CASE(256) AZB_CASE(512)
    AZB_CASE(1024) AZB_CASE(48) AZB_CASE(96) AZB_CASE(192) AZB_CASE(384) AZB_CASE(768)
    AZB_CASE(1536)
#undef AZB_CASE
    default:
      return fail(AZB_ERR_UNSUPPORTED, "FFT length " + std::to_string(L) + " not built");
  }
  return AZB_OK;
}

template <typename T, int DIM>
int launch_row(int L, RowArgs a, long nsamples, hipStream_t st, int pdiv = 1) {
  int P = row_p(L);
  if (pdiv == 2 && P >= 2) P /= 2;
  if (a.nlines % (2 * P) != 0) return fail(AZB_ERR_UNSUPPORTED, "line count not a multiple of the line-pair tile");
  if (DIM == 3 && L == 1536) return fail(AZB_ERR_UNSUPPORTED, "3D row length 1536 not built");
  const int r = launch_row_fused((int)sizeof(T), DIM, L, a, nsamples, st, pdiv);
  if (r == 2) return fail(AZB_ERR_UNSUPPORTED, "line count not a multiple of the line-pair tile");
  if (r != 0) return fail(AZB_ERR_UNSUPPORTED, "row length " + std::to_string(L) + " not built");
  return AZB_OK;
}

template <typename T, int L, int DIR>
void launch_plain_row_L(PlainRowArgs a, hipStream_t st) {
  constexpr int P = (L % 3 == 0) ? ((1536 / L) < 1 ? 1 : 1536 / L) : ((1024 / L) < 1 ? 1 : 1024 / L);
  constexpr int NT = P * L / LenPlan<L>::E;
  const long blocks = (a.nlines + 2 * P - 1) / (2 * P);
  hipLaunchKernelGGL((plain_row_kernel<T, L, P, DIR>), dim3((unsigned)blocks), dim3(NT), 0, st, a);
}

template <typename T, int DIR>
int launch_plain_row(int L, PlainRowArgs a, hipStream_t st) {
  switch (L) {
#define AZB_CASE(LL) \
  case LL: launch_plain_row_L<T, LL, DIR>(a, st); break;
    AZB_CASE(16) AZB_CASE(32) AZB_CASE(64) AZB_CASE(128) AZB_CASE(256) AZB_CASE(512)
    AZB_CASE(1024) AZB_CASE(48) AZB_CASE(96) AZB_CASE(192) AZB_CASE(384) AZB_CASE(768)
    AZB_CASE(1536)
#undef AZB_CASE
    default:
      return fail(AZB_ERR_UNSUPPORTED, "FFT length " + std::to_string(L) + " not built");
  }
  return AZB_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------
struct azb_solver {
  int dim = 0, N = 0, M = 0, nh = 0, nvars = 0, prec = 0;
  int S = 1;  // samples solved together (same grid), stacked outermost
  long nhp = 0;  // last-axis stride of the intermediates: N/2+1 rounded up to 16
  azb_sv_params visc{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  size_t tsize = 8;  // bytes of one real at working precision
  void* X = nullptr;  // T1 / S2
  void* Y = nullptr;  // T2 / B_hat
  void* Z = nullptr;  // S1
  double2* u1 = nullptr;
  double2* u2 = nullptr;
  DevState* st = nullptr;
  void* tw = nullptr;
  size_t bytes = 0;
  long nmodes = 0;  // modes per component on the N-grid
  double loop_C = 0.2;  // CFL number of the device-driven loop
  bool loop_begun = false;  // loop_C fixed by azb_loop_begin / azb_loop_begin_sample
  // tuning: column tile divisor per kernel family, line-pair divisor of the row kernel
  int tsdiv[AZB_PROF_FAMILIES] = {1, 1, 1, 1, 1, 1};
  int pdiv = 1;
  // radix plan per column-pass family (ColPlan index; the row and k-space entries are 0)
  int plan[AZB_PROF_FAMILIES] = {0, 0, 0, 0, 0, 0};
  DevState* probe_host = nullptr;  // pinned mirror of the device state (non-blocking probes)
  hipEvent_t probe_ev = nullptr;
  // per-kernel-family hipEvent timing (off unless enabled)
  bool prof = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_ev;
  std::vector<hipEvent_t> ev_pool;
  double prof_ms[AZB_PROF_FAMILIES] = {};
  long prof_n[AZB_PROF_FAMILIES] = {};
};

namespace {

hipEvent_t pool_event(azb_solver* s) {
  if (!s->ev_pool.empty()) {
    hipEvent_t e = s->ev_pool.back();
    s->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// RAII bracket of one kernel family launch on the solver's stream
struct ProfScope {
  azb_solver* s;
  int fam;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(azb_solver* s_, int fam_) : s(s_), fam(fam_) {
    if (s->prof) {
      a = pool_event(s);
      b = pool_event(s);
      (void)hipEventRecord(a, s->stream);
    }
  }
  ~ProfScope() {
    if (s->prof) {
      (void)hipEventRecord(b, s->stream);
      s->prof_ev.push_back({fam, {a, b}});
    }
  }
};

template <typename T>
int enqueue_rhs_T(azb_solver* s, const double2* u_in, const double2* u0, double2* out, int kmode,
                  int slot, double C) {
  const int N = s->N, M = s->M, nh = s->nh;
  const long S = s->S;
  const long hp = s->nhp;  // padded last-axis stride of every intermediate
  hipStream_t st = s->stream;
  int rc;
  ColArgs c{};
  c.nsmall = N;
  c.tw = s->tw;
  c.st = s->st;
  c.nk = nh;
  RowArgs r{};
  r.tw = s->tw;
  r.st = s->st;
  r.slot = slot;
  r.ls = hp;
  r.norm = 1.0 / (std::pow(std::sqrt((double)N), s->dim) * std::pow(std::sqrt((double)M), s->dim));
  if (s->dim == 3) {
    // K1: inverse along axis 0, N -> M rows; state [c][i][j][k] -> X[c][i'][j][hp]
    c.in = u_in; c.out = s->X; c.nq = N;
    c.in_qs = nh; c.out_qs = hp;
    c.in_rs = (long)N * nh; c.out_rs = (long)N * hp;
    c.in_bs = (long)N * N * nh; c.out_bs = (long)M * N * hp;
    c.bps = 3;
    {
      ProfScope p(s, AZB_PROF_INV_AX0);
      if ((rc = launch_col<double, T, T, +1, COL_PAD_IN>(M, c, 3L * S, st, s->tsdiv[AZB_PROF_INV_AX0], s->plan[AZB_PROF_INV_AX0]))) return rc;
    }
    // K2: inverse along axis 1; X[c][i'][j][hp] -> Y[c][i'][j'][hp]
    c.in = s->X; c.out = s->Y; c.nq = 1;
    c.in_qs = 0; c.out_qs = 0;
    c.in_rs = hp; c.out_rs = hp;
    c.in_bs = (long)N * hp; c.out_bs = (long)M * hp;
    c.bps = 3 * M;
    {
      ProfScope p(s, AZB_PROF_INV_AX1);
      if ((rc = launch_col<T, T, T, +1, COL_PAD_IN>(M, c, 3L * M * S, st, s->tsdiv[AZB_PROF_INV_AX1], s->plan[AZB_PROF_INV_AX1]))) return rc;
    }
    // K3: last axis c2r, products, max, r2c; Y[c][line][hp] -> Z[m][line][hp]
    r.in = s->Y; r.out = s->Z; r.nlines = (long)M * M;
    r.in_ss = 3L * M * M * hp; r.out_ss = 6L * M * M * hp;
    {
      ProfScope p(s, AZB_PROF_ROW);
      if ((rc = launch_row<T, 3>(M, r, S, st, s->pdiv))) return rc;
    }
    // K4: forward along axis 1, M -> N rows; Z[m][i'][j'][hp] -> X[m][i'][j][hp]
    c.in = s->Z; c.out = s->X; c.nq = 1;
    c.in_qs = 0; c.out_qs = 0;
    c.in_rs = hp; c.out_rs = hp;
    c.in_bs = (long)M * hp; c.out_bs = (long)N * hp;
    c.bps = 6 * M;
    {
      ProfScope p(s, AZB_PROF_FWD_AX1);
      if ((rc = launch_col<T, T, T, -1, COL_TRUNC_OUT>(M, c, 6L * M * S, st, s->tsdiv[AZB_PROF_FWD_AX1], s->plan[AZB_PROF_FWD_AX1]))) return rc;
    }
    // K5: forward along axis 0; X[m][i'][j][hp] -> Y[m][i][j][hp] (B_hat on the N-grid)
    c.in = s->X; c.out = s->Y; c.nq = N;
    c.in_qs = hp; c.out_qs = hp;
    c.in_rs = (long)N * hp; c.out_rs = (long)N * hp;
    c.in_bs = (long)M * N * hp; c.out_bs = (long)N * N * hp;
    c.bps = 6;
    {
      ProfScope p(s, AZB_PROF_FWD_AX0);
      if ((rc = launch_col<T, T, T, -1, COL_TRUNC_OUT>(M, c, 6L * S, st, s->tsdiv[AZB_PROF_FWD_AX0], s->plan[AZB_PROF_FWD_AX0]))) return rc;
    }
  } else {
    // K1: inverse along axis 0; state [c][i][k] -> X[c][i'][hp]
    c.in = u_in; c.out = s->X; c.nq = 1;
    c.in_qs = 0; c.out_qs = 0;
    c.in_rs = nh; c.out_rs = hp;
    c.in_bs = (long)N * nh; c.out_bs = (long)M * hp;
    c.bps = 2;
    {
      ProfScope p(s, AZB_PROF_INV_AX0);
      if ((rc = launch_col<double, T, T, +1, COL_PAD_IN>(M, c, 2L * S, st, s->tsdiv[AZB_PROF_INV_AX0], s->plan[AZB_PROF_INV_AX0]))) return rc;
    }
    // K2: last axis; X[c][i'][hp] -> Z[m][i'][hp]
    r.in = s->X; r.out = s->Z; r.nlines = M;
    r.in_ss = 2L * M * hp; r.out_ss = 3L * M * hp;
    {
      ProfScope p(s, AZB_PROF_ROW);
      if ((rc = launch_row<T, 2>(M, r, S, st, s->pdiv))) return rc;
    }
    // K3: forward along axis 0; Z[m][i'][hp] -> Y[m][i][hp]
    c.in = s->Z; c.out = s->Y; c.nq = 1;
    c.in_qs = 0; c.out_qs = 0;
    c.in_rs = hp; c.out_rs = hp;
    c.in_bs = (long)M * hp; c.out_bs = (long)N * hp;
    c.bps = 3;
    {
      ProfScope p(s, AZB_PROF_FWD_AX0);
      if ((rc = launch_col<T, T, T, -1, COL_TRUNC_OUT>(M, c, 3L * S, st, s->tsdiv[AZB_PROF_FWD_AX0], s->plan[AZB_PROF_FWD_AX0]))) return rc;
    }
  }
  const double Md = std::pow((double)M, s->dim);
  KArgs k{};
  // 2D: the max|u| reduction runs inside the k-space kernel (a launch less per RHS
  // where launches are short); 3D keeps it in its own one-workgroup kernel rather
  // than have each of a million waves re-read the bins at stage 1
  k.fold = s->dim == 2 ? 1 : 0;
  k.slot = slot;
  k.Md = Md;
  k.C = C;
  k.Nd = (double)N;
  if (!k.fold)
    hipLaunchKernelGGL(finalize_umax, dim3((unsigned)S), dim3(64), 0, st, s->st, slot, kmode == K_RK1 ? 1 : 0,
                       Md, C, (double)N);
  k.bh = s->Y;
  k.u_in = u_in;
  k.u0 = u0;
  k.out = out;
  k.nmodes = s->nmodes;
  k.n = N;
  k.nh = nh;
  k.nhp = (int)hp;
  k.bstride = s->dim == 3 ? (long)N * N * hp : (long)N * hp;
  k.eps = s->visc.eps;
  k.s2 = 2 * s->visc.s;
  k.k0 = s->visc.k0;
  k.mode = kmode;
  k.st = s->st;
  const unsigned blocks = (unsigned)((s->nmodes + 255) / 256);
  {
    ProfScope p(s, AZB_PROF_KSPACE);
    launch_kspace((int)sizeof(T), s->dim, k, blocks, (unsigned)S, st);
  }
  HIPCHK(hipGetLastError());
  return AZB_OK;
}

int enqueue_rhs(azb_solver* s, const double2* u_in, const double2* u0, double2* out, int kmode,
                int slot, double C) {
  if (s->prec == AZB_FP64) return enqueue_rhs_T<double>(s, u_in, u0, out, kmode, slot, C);
  return enqueue_rhs_T<float>(s, u_in, u0, out, kmode, slot, C);
}

// the three stages of SSP_RK3::integrate (ssp_rk3.hpp:54-62), stage updates fused
int enqueue_rk3_stages(azb_solver* s, double2* u, double C) {
  int rc;
  if ((rc = enqueue_rhs(s, u, nullptr, s->u1, K_RK1, 0, C))) return rc;   // u1 = u + dt L(u)
  if ((rc = enqueue_rhs(s, s->u1, u, s->u2, K_RK2, 1, C))) return rc;     // u2 = 3/4 u + 1/4 (u1 + dt L(u1))
  if ((rc = enqueue_rhs(s, s->u2, u, u, K_RK3, 2, C))) return rc;         // u = 2/3 (u2 + dt L(u2)) + 1/3 u
  return AZB_OK;
}

int check_solver_args(int dim, int N, int M, int n_vars, int precision, int nsamples) {
  if (dim != 2 && dim != 3) return fail(AZB_ERR_ARG, "dim must be 2 or 3");
  if (n_vars != dim) return fail(AZB_ERR_UNSUPPORTED, "n_vars must equal dim (no tracer)");
  if (N < 8 || N % 4 != 0) return fail(AZB_ERR_ARG, "N must be a multiple of 4");
  if (nsamples < 1 || nsamples > 65535) return fail(AZB_ERR_ARG, "nsamples must be in [1, 65535]");
  if (M == 0) M = 3 * N / 2;
  if (M != 3 * N / 2) return fail(AZB_ERR_UNSUPPORTED, "only the default padding M = 3N/2 is built");
  if (M != 96 && M != 192 && M != 384 && M != 768 && M != 1536)
    return fail(AZB_ERR_UNSUPPORTED, "padded size " + std::to_string(M) + " not built");
  if (dim == 3 && M == 1536) return fail(AZB_ERR_UNSUPPORTED, "3D N = 1024 is not built");
  if (precision != AZB_FP64 && precision != AZB_FP32_MIXED)
    return fail(AZB_ERR_ARG, "unknown precision mode");
  return AZB_OK;
}

// bytes of the intermediates X, Y, Z and of one RK stage buffer for S samples
struct Layout {
  size_t xs, ys, zs, us;
};

// fill a solver's grid fields (arguments already checked); no allocation
Layout setup_solver(azb_solver* s, int dim, int N, int precision, azb_sv_params visc, int nsamples) {
  const int M = 3 * N / 2;
  s->dim = dim;
  s->N = N;
  s->M = M;
  s->S = nsamples;
  s->nh = N / 2 + 1;
  s->nvars = dim;
  s->prec = precision;
  s->visc = visc;
  // 3D: the 12-values-per-thread row kernel with 192-thread workgroups
  s->pdiv = dim == 3 ? 2 : 1;
  s->tsize = precision == AZB_FP64 ? 8 : 4;
  const size_t cs = 2 * s->tsize;  // complex at working precision
  const long nh = s->nh;
  s->nhp = (nh + 15) / 16 * 16;
  const long hp = s->nhp;
  const size_t S = nsamples;
  Layout l{};
  if (dim == 3) {
    s->nmodes = (long)N * N * nh;
    l.xs = S * cs * 6 * (size_t)M * N * hp;                                   // T1 (3) / S2 (6)
    l.ys = S * cs * std::max<size_t>(3 * (size_t)M * M * hp, 6 * (size_t)N * N * hp);  // T2 / B_hat
    l.zs = S * cs * 6 * (size_t)M * M * hp;                                   // S1
  } else {
    s->nmodes = (long)N * nh;
    l.xs = S * cs * 2 * (size_t)M * hp;
    l.ys = S * cs * 3 * (size_t)N * hp;
    l.zs = S * cs * 3 * (size_t)M * hp;
  }
  l.us = S * sizeof(double2) * (size_t)dim * s->nmodes;
  return l;
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// ---------------------------------------------------------------------------
// Scratch of the no-allocation RHS (azb_dudt_noalloc): the twiddle tables of every
// built padded length (FP64, then FP32) at the start, then the intermediates of one
// solve.  Each GPU holds its own copy of the array for the life of the process.
// ---------------------------------------------------------------------------
constexpr int kTwLens[5] = {96, 192, 384, 768, 1536};
constexpr size_t kScratchTwBytes = size_t(128) << 10;
constexpr size_t kScratchBytes = size_t(96) << 20;

size_t scratch_tw_offset(int precision, int M) {
  size_t off = 0;
  if (precision != AZB_FP64)
    for (int L : kTwLens) off += sizeof(cx<double>) * twiddle_entries<double>(L);
  for (int L : kTwLens) {
    if (L == M) break;
    off += precision == AZB_FP64 ? sizeof(cx<double>) * twiddle_entries<double>(L)
                                 : sizeof(cx<float>) * twiddle_entries<float>(L);
  }
  return off;
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;  // device address of `scratch`, per device

// the current device's address of the scratch array; on a device's first use it is
// looked up and the twiddle tables are written (the name `scratch` is a host-side
// stand-in: only this address may reach a kernel or a copy)
int scratch_base(unsigned char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    unsigned char* base = (unsigned char*)p;
    for (int L : kTwLens) {
      const auto hd = host_tables<double>(L);
      HIPCHK(hipMemcpy(base + scratch_tw_offset(AZB_FP64, L), hd.data(), sizeof(cx<double>) * hd.size(),
                       hipMemcpyHostToDevice));
      const auto hf = host_tables<float>(L);
      HIPCHK(hipMemcpy(base + scratch_tw_offset(AZB_FP32_MIXED, L), hf.data(), sizeof(cx<float>) * hf.size(),
                       hipMemcpyHostToDevice));
    }
    g_scratch_ptr[dev] = base;
  }
  *out = g_scratch_ptr[dev];
  return AZB_OK;
}

static_assert(sizeof(cx<double>) * (96 + 192 + 384 + 768 + 1536) +
                      sizeof(cx<float>) * (96 + 192 + 384 + kTw768F32End + 1536) <=
                  kScratchTwBytes,
              "twiddle tables must fit their region of the scratch array");

}  // namespace

extern "C" {

const char* azb_last_error(void) { return g_err.c_str(); }

int azb_create_batch(int dim, int N, int M, int n_vars, int precision, azb_sv_params visc, int nsamples,
                     void* stream, azb_solver** out) {
  if (!out) return fail(AZB_ERR_ARG, "out is NULL");
  *out = nullptr;
  int rc = check_solver_args(dim, N, M, n_vars, precision, nsamples);
  if (rc) return rc;
  azb_solver* s = new azb_solver();
  const Layout lay = setup_solver(s, dim, N, precision, visc, nsamples);
  const size_t xs = lay.xs, ys = lay.ys, zs = lay.zs, us = lay.us, S = nsamples;
  auto cleanup = [&](int code, const std::string& m) {
    azb_destroy(s);
    return fail(code, m);
  };
  if (hipMalloc(&s->X, xs) != hipSuccess || hipMalloc(&s->Y, ys) != hipSuccess ||
      hipMalloc(&s->Z, zs) != hipSuccess || hipMalloc((void**)&s->u1, us) != hipSuccess ||
      hipMalloc((void**)&s->u2, us) != hipSuccess ||
      hipMalloc((void**)&s->st, S * sizeof(DevState)) != hipSuccess)
    return cleanup(AZB_ERR_ALLOC, "device allocation of the solver workspace failed");
  s->bytes = xs + ys + zs + 2 * us + S * sizeof(DevState);
  rc = precision == AZB_FP64 ? make_twiddles<double>(s->M, &s->tw) : make_twiddles<float>(s->M, &s->tw);
  if (rc) {
    std::string m = g_err;
    return cleanup(rc, m);
  }
  s->stream = (hipStream_t)stream;  // NULL: the default (null) stream
  std::vector<DevState> h(S);
  if (hipMemcpy(s->st, h.data(), S * sizeof(DevState), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(AZB_ERR_HIP, "hipMemcpy of the device state failed");
  *out = s;
  return AZB_OK;
}

int azb_create(int dim, int N, int M, int n_vars, int precision, azb_sv_params visc, void* stream,
               azb_solver** out) {
  return azb_create_batch(dim, N, M, n_vars, precision, visc, 1, stream, out);
}

int azb_num_samples(const azb_solver* s) { return s ? s->S : 0; }

size_t azb_bytes_estimate(int dim, int N, int precision, int nsamples) {
  const size_t cs = precision == AZB_FP64 ? 16 : 8;
  const size_t M = 3 * (size_t)N / 2, nh = N / 2 + 1, hp = (nh + 15) / 16 * 16, S = nsamples;
  size_t ws, us;
  if (dim == 3) {
    ws = cs * (6 * M * N * hp + std::max(3 * M * M * hp, 6 * (size_t)N * N * hp) + 6 * M * M * hp);
    us = sizeof(double2) * 3 * (size_t)N * N * nh;
  } else {
    ws = cs * (2 * M * hp + 3 * (size_t)N * hp + 3 * M * hp);
    us = sizeof(double2) * 2 * (size_t)N * nh;
  }
  return S * (ws + 2 * us + sizeof(DevState));
}

int azb_destroy(azb_solver* s) {
  if (!s) return AZB_OK;
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(s->X);
  (void)hipFree(s->Y);
  (void)hipFree(s->Z);
  (void)hipFree(s->u1);
  (void)hipFree(s->u2);
  (void)hipFree(s->st);
  (void)hipFree(s->tw);
  if (s->probe_host) (void)hipHostFree(s->probe_host);
  if (s->probe_ev) (void)hipEventDestroy(s->probe_ev);
  for (auto& e : s->prof_ev) {
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  for (auto e : s->ev_pool) (void)hipEventDestroy(e);
  if (s->own_stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return AZB_OK;
}

size_t azb_workspace_bytes(const azb_solver* s) { return s ? s->bytes : 0; }
void* azb_stream(const azb_solver* s) { return s ? (void*)s->stream : nullptr; }

int azb_dudt(azb_solver* s, const void* u_hat, void* dudt_hat) {
  if (!s || !u_hat || !dudt_hat) return fail(AZB_ERR_ARG, "NULL argument");
  hipLaunchKernelGGL(step_prologue_fixed, dim3(s->S), dim3(256), 0, s->stream, s->st, 0.0);
  return enqueue_rhs(s, (const double2*)u_hat, nullptr, (double2*)dudt_hat, K_DUDT, 3, 0.0);
}

int azb_dudt_noalloc(int dim, int N, int precision, azb_sv_params visc, const void* u_hat, void* dudt_hat,
                     void* stream) {
  if (!u_hat || !dudt_hat) return fail(AZB_ERR_ARG, "NULL argument");
  int rc = check_solver_args(dim, N, 0, dim, precision, 1);
  if (rc) return rc;
  azb_solver s;  // grid fields only: its buffers are carved out of the scratch array below
  const Layout l = setup_solver(&s, dim, N, precision, visc, 1);
  const size_t xs = round256(l.xs), ys = round256(l.ys), zs = round256(l.zs);
  const size_t need = xs + ys + zs + round256(sizeof(DevState));
  if (need > kScratchBytes - kScratchTwBytes)
    return fail(AZB_ERR_UNSUPPORTED, std::to_string(dim) + "D N = " + std::to_string(N) + " needs " +
                                         std::to_string(need) + " bytes of scratch; the no-allocation RHS has " +
                                         std::to_string(kScratchBytes - kScratchTwBytes) +
                                         " (2D N <= 1024, 3D N <= 64)");
  unsigned char* base = nullptr;
  if ((rc = scratch_base(&base))) return rc;
  s.tw = base + scratch_tw_offset(precision, s.M);
  unsigned char* w = base + kScratchTwBytes;
  s.X = w;
  s.Y = w + xs;
  s.Z = w + xs + ys;
  s.st = (DevState*)(w + xs + ys + zs);
  s.stream = (hipStream_t)stream;
  // clears the max|u| bins the previous call on this GPU left in the array
  hipLaunchKernelGGL(step_prologue_fixed, dim3(1), dim3(256), 0, s.stream, s.st, 0.0);
  return enqueue_rhs(&s, (const double2*)u_hat, nullptr, (double2*)dudt_hat, K_DUDT, 3, 0.0);
}

int azb_u_max(azb_solver* s, double* u_max_host) {
  if (!s || !u_max_host) return fail(AZB_ERR_ARG, "NULL argument");
  HIPCHK(hipMemcpyAsync(u_max_host, &s->st->u_max, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return AZB_OK;
}

int azb_dt(azb_solver* s, double C, double* dt_host) {
  double um = 0;
  int rc = azb_u_max(s, &um);
  if (rc) return rc;
  if (!dt_host) return fail(AZB_ERR_ARG, "NULL argument");
  *dt_host = C / (s->N * um);
  return AZB_OK;
}

int azb_rk3_step(azb_solver* s, void* u_hat, double max_dt, double C, double* dt_host) {
  if (!s || !u_hat) return fail(AZB_ERR_ARG, "NULL argument");
  hipLaunchKernelGGL(step_prologue_fixed, dim3(s->S), dim3(256), 0, s->stream, s->st, max_dt);
  int rc = enqueue_rk3_stages(s, (double2*)u_hat, C);
  if (rc) return rc;
  if (dt_host) {
    HIPCHK(hipMemcpyAsync(dt_host, &s->st->dt, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  return AZB_OK;
}

int azb_loop_begin(azb_solver* s, double t_end, double C, const double* times0_host) {
  if (!s) return fail(AZB_ERR_ARG, "NULL solver");
  const double pi = 3.14159265358979323846;
  std::vector<DevState> h(s->S);
  for (int i = 0; i < s->S; ++i) {
    h[i].time = times0_host ? times0_host[i] : 0.0;
    h[i].t_end = t_end;
    h[i].max_dt_visc = C * 2. / (s->visc.eps * ((pi * s->N) * (pi * s->N)));
    h[i].done = 0;
    h[i].steps = 0;
  }
  s->loop_C = C;
  s->loop_begun = true;
  HIPCHK(hipMemcpyAsync(s->st, h.data(), s->S * sizeof(DevState), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return AZB_OK;
}

int azb_loop_begin_sample(azb_solver* s, int sample, double t_end, double C, double time0) {
  if (!s) return fail(AZB_ERR_ARG, "NULL solver");
  if (sample < 0 || sample >= s->S)
    return fail(AZB_ERR_ARG, "sample " + std::to_string(sample) + " outside [0, " + std::to_string(s->S) + ")");
  if (C != s->loop_C && s->loop_begun)
    return fail(AZB_ERR_ARG, "every sample of a solver shares one CFL number");
  const double pi = 3.14159265358979323846;
  s->loop_C = C;
  s->loop_begun = true;
  hipLaunchKernelGGL(loop_set_sample, dim3(1), dim3(64), 0, s->stream, s->st, sample, time0, t_end,
                     C * 2. / (s->visc.eps * ((pi * s->N) * (pi * s->N))));
  HIPCHK(hipGetLastError());
  return AZB_OK;
}

int azb_enqueue_steps(azb_solver* s, void* u_hat, int nsteps) {
  if (!s || !u_hat) return fail(AZB_ERR_ARG, "NULL argument");
  for (int i = 0; i < nsteps; ++i) {
    hipLaunchKernelGGL(step_prologue_loop, dim3(s->S), dim3(256), 0, s->stream, s->st);
    int rc = enqueue_rk3_stages(s, (double2*)u_hat, s->loop_C);
    if (rc) return rc;
    hipLaunchKernelGGL(step_epilogue_loop, dim3(s->S), dim3(64), 0, s->stream, s->st);
  }
  HIPCHK(hipGetLastError());
  return AZB_OK;
}

int azb_loop_read(azb_solver* s, double* times_host, int* steps_host) {
  if (!s) return fail(AZB_ERR_ARG, "NULL solver");
  std::vector<DevState> h(s->S);
  HIPCHK(hipMemcpyAsync(h.data(), s->st, s->S * sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int i = 0; i < s->S; ++i) {
    if (times_host) times_host[i] = h[i].time;
    if (steps_host) steps_host[i] = h[i].steps;
  }
  return AZB_OK;
}

int azb_loop_probe(azb_solver* s) {
  if (!s) return fail(AZB_ERR_ARG, "NULL solver");
  if (!s->probe_host) {
    HIPCHK(hipHostMalloc((void**)&s->probe_host, s->S * sizeof(DevState), hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&s->probe_ev, hipEventDisableTiming));
  }
  HIPCHK(hipMemcpyAsync(s->probe_host, s->st, s->S * sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipEventRecord(s->probe_ev, s->stream));
  return AZB_OK;
}

int azb_loop_probe_result(azb_solver* s, int* ready, double* times_host, int* steps_host) {
  if (!s || !ready) return fail(AZB_ERR_ARG, "NULL argument");
  if (!s->probe_ev) return fail(AZB_ERR_ARG, "no probe queued");
  const hipError_t e = hipEventQuery(s->probe_ev);
  if (e == hipErrorNotReady) {
    *ready = 0;
    return AZB_OK;
  }
  if (e != hipSuccess) return fail(AZB_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
  *ready = 1;
  for (int i = 0; i < s->S; ++i) {
    if (times_host) times_host[i] = s->probe_host[i].time;
    if (steps_host) steps_host[i] = s->probe_host[i].steps;
  }
  return AZB_OK;
}

int azb_memcpy_h2d_async(void* dst, const void* src_host, size_t bytes, void* stream) {
  HIPCHK(hipMemcpyAsync(dst, src_host, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return AZB_OK;
}

int azb_memcpy_d2h_async(void* dst_host, const void* src, size_t bytes, void* stream) {
  HIPCHK(hipMemcpyAsync(dst_host, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return AZB_OK;
}

int azb_simulate_until(azb_solver* s, void* u_hat, double t_end, double C, double* time_inout,
                       int max_steps, int chunk, int* steps_out) {
  if (!s || !u_hat || !time_inout) return fail(AZB_ERR_ARG, "NULL argument");
  if (chunk <= 0) chunk = 8;
  std::vector<double> t0(s->S, *time_inout), times(s->S);
  std::vector<int> steps(s->S);
  int rc = azb_loop_begin(s, t_end, C, t0.data());
  if (rc) return rc;
  int taken = 0;
  while (true) {
    int n = chunk;
    if (max_steps > 0 && taken + n > max_steps) n = max_steps - taken;
    if ((rc = azb_enqueue_steps(s, u_hat, n))) return rc;
    taken += n;
    if ((rc = azb_loop_read(s, times.data(), steps.data()))) return rc;
    bool all_done = true;
    for (double tt : times) all_done = all_done && !(tt < t_end);
    if (all_done) break;
    if (max_steps > 0 && taken >= max_steps) break;
  }
  int mx = 0;
  double tmin = t_end;
  for (int i = 0; i < s->S; ++i) {
    mx = std::max(mx, steps[i]);
    tmin = std::min(tmin, times[i]);
  }
  if (steps_out) *steps_out = mx;
  *time_inout = tmin;
  return AZB_OK;
}

// ---- hipGraph capture of enqueued work -------------------------------------
struct azb_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int azb_capture_begin(void* stream) {
  if (!stream) return fail(AZB_ERR_ARG, "capture needs a non-null stream");
  HIPCHK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return AZB_OK;
}

int azb_capture_end(void* stream, azb_graph** out) {
  if (!stream || !out) return fail(AZB_ERR_ARG, "NULL argument");
  azb_graph* g = new azb_graph();
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
  if (e != hipSuccess) {
    delete g;
    return fail(AZB_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    delete g;
    return fail(AZB_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  *out = g;
  return AZB_OK;
}

int azb_graph_launch(azb_graph* g, void* stream) {
  if (!g) return fail(AZB_ERR_ARG, "NULL graph");
  HIPCHK(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return AZB_OK;
}

int azb_graph_destroy(azb_graph* g) {
  if (!g) return AZB_OK;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return AZB_OK;
}

int azb_stream_create(void** stream) {
  if (!stream) return fail(AZB_ERR_ARG, "NULL argument");
  hipStream_t st = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *stream = (void*)st;
  return AZB_OK;
}

int azb_stream_destroy(void* stream) {
  HIPCHK(hipStreamDestroy((hipStream_t)stream));
  return AZB_OK;
}

int azb_stream_synchronize(void* stream) {
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return AZB_OK;
}

int azb_stream_wait(void* waiter, void* waitee) {
  if (waiter == waitee) return AZB_OK;
  hipEvent_t e = nullptr;
  HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipError_t r = hipEventRecord(e, (hipStream_t)waitee);
  if (r == hipSuccess) r = hipStreamWaitEvent((hipStream_t)waiter, e, 0);
  // the wait holds what it needs: the event may be released at once
  (void)hipEventDestroy(e);
  if (r != hipSuccess) return fail(AZB_ERR_HIP, std::string("azb_stream_wait: ") + hipGetErrorString(r));
  return AZB_OK;
}

int azb_profile_enable(azb_solver* s, int on) {
  if (!s) return fail(AZB_ERR_ARG, "NULL solver");
  s->prof = on != 0;
  return AZB_OK;
}

int azb_profile_read(azb_solver* s, double* ms_host, long* launches_host, int n) {
  if (!s || !ms_host || !launches_host) return fail(AZB_ERR_ARG, "NULL argument");
  HIPCHK(hipStreamSynchronize(s->stream));
  for (auto& e : s->prof_ev) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
    s->prof_ms[e.first] += ms;
    s->prof_n[e.first] += 1;
    s->ev_pool.push_back(e.second.first);
    s->ev_pool.push_back(e.second.second);
  }
  s->prof_ev.clear();
  for (int i = 0; i < n && i < AZB_PROF_FAMILIES; ++i) {
    ms_host[i] = s->prof_ms[i];
    launches_host[i] = s->prof_n[i];
  }
  return AZB_OK;
}

int azb_profile_reset(azb_solver* s) {
  if (!s) return fail(AZB_ERR_ARG, "NULL solver");
  HIPCHK(hipStreamSynchronize(s->stream));
  for (auto& e : s->prof_ev) {
    s->ev_pool.push_back(e.second.first);
    s->ev_pool.push_back(e.second.second);
  }
  s->prof_ev.clear();
  for (int i = 0; i < AZB_PROF_FAMILIES; ++i) {
    s->prof_ms[i] = 0;
    s->prof_n[i] = 0;
  }
  return AZB_OK;
}

int azb_set_tuning(azb_solver* s, const int* tsdiv, int pdiv) {
  if (!s || !tsdiv) return fail(AZB_ERR_ARG, "NULL argument");
  for (int i = 0; i < AZB_PROF_FAMILIES; ++i) {
    const int v = tsdiv[i];
    if (v != 1 && v != 2 && v != 4) return fail(AZB_ERR_ARG, "tile divisors must be 1, 2 or 4");
    s->tsdiv[i] = v;
  }
  if (pdiv != 1 && pdiv != 2) return fail(AZB_ERR_ARG, "row divisor must be 1 or 2");
  const int P = row_p(s->M);
  const long lines = s->dim == 3 ? (long)s->M * s->M : (long)s->M;
  s->pdiv = (pdiv == 2 && P >= 2 && lines % P == 0) ? 2 : 1;
  return AZB_OK;
}

int azb_get_tuning(const azb_solver* s, int* tsdiv, int* pdiv) {
  if (!s || !tsdiv || !pdiv) return fail(AZB_ERR_ARG, "NULL argument");
  for (int i = 0; i < AZB_PROF_FAMILIES; ++i) tsdiv[i] = s->tsdiv[i];
  *pdiv = s->pdiv;
  return AZB_OK;
}

int azb_autotune(azb_solver* s, const void* u_hat, int reps) {
  if (!s || !u_hat) return fail(AZB_ERR_ARG, "NULL argument");
  if (reps < 1) reps = 3;
  const bool prof_was = s->prof;
  const int cands[3] = {1, 2, 4};
  double best[AZB_PROF_FAMILIES];
  int pick[AZB_PROF_FAMILIES];
  for (int f = 0; f < AZB_PROF_FAMILIES; ++f) {
    best[f] = 1e300;
    pick[f] = 1;
  }
  double best_row = 1e300;
  int pick_row = 1;
  int rc = AZB_OK;
  for (int ci = 0; ci < 3; ++ci) {
    const int c = cands[ci];
    int td[AZB_PROF_FAMILIES];
    for (int f = 0; f < AZB_PROF_FAMILIES; ++f) td[f] = c;
    if ((rc = azb_set_tuning(s, td, c >= 2 ? 2 : 1))) break;
    const bool row_variant_ok = (c == 1) || s->pdiv == 2;
    // warm-up, then timed evaluations into the solver's own stage buffer
    if ((rc = azb_dudt(s, u_hat, s->u1))) break;
    if ((rc = azb_profile_reset(s))) break;
    s->prof = true;
    for (int r = 0; r < reps && rc == AZB_OK; ++r) rc = azb_dudt(s, u_hat, s->u1);
    s->prof = false;
    if (rc) break;
    double ms[AZB_PROF_FAMILIES];
    long n[AZB_PROF_FAMILIES];
    if ((rc = azb_profile_read(s, ms, n, AZB_PROF_FAMILIES))) break;
    for (int f = 0; f < AZB_PROF_FAMILIES; ++f) {
      if (n[f] == 0) continue;
      const double avg = ms[f] / n[f];
      if (f == AZB_PROF_ROW) {
        if (row_variant_ok && avg < best_row) {
          best_row = avg;
          pick_row = c >= 2 ? 2 : 1;
        }
      } else if (f != AZB_PROF_KSPACE && avg < best[f]) {
        best[f] = avg;
        pick[f] = c;
      }
    }
  }
  s->prof = prof_was;
  (void)azb_profile_reset(s);
  if (rc) return rc;
  return azb_set_tuning(s, pick, pick_row);
}

int azb_num_plans(const azb_solver* s) { return s ? col_plan_count(s->M) : 0; }

int azb_set_plans(azb_solver* s, const int* plans) {
  if (!s || !plans) return fail(AZB_ERR_ARG, "NULL argument");
  const int n = col_plan_count(s->M);
  for (int f = 0; f < AZB_PROF_FAMILIES; ++f) {
    const bool col = f != AZB_PROF_ROW && f != AZB_PROF_KSPACE;
    if (plans[f] < 0 || plans[f] >= (col ? n : 1))
      return fail(AZB_ERR_ARG, "radix plan " + std::to_string(plans[f]) + " of family " + std::to_string(f) +
                                   " outside [0, " + std::to_string(col ? n : 1) + ")");
  }
  for (int f = 0; f < AZB_PROF_FAMILIES; ++f) s->plan[f] = plans[f];
  return AZB_OK;
}

int azb_get_plans(const azb_solver* s, int* plans) {
  if (!s || !plans) return fail(AZB_ERR_ARG, "NULL argument");
  for (int f = 0; f < AZB_PROF_FAMILIES; ++f) plans[f] = s->plan[f];
  return AZB_OK;
}

int azb_autotune_plans(azb_solver* s, const void* u_hat, int reps) {
  if (!s || !u_hat) return fail(AZB_ERR_ARG, "NULL argument");
  if (reps < 1) reps = 3;
  const int n = col_plan_count(s->M);
  if (n == 1) return AZB_OK;
  const bool prof_was = s->prof;
  int keep[AZB_PROF_FAMILIES], pick[AZB_PROF_FAMILIES];
  double best[AZB_PROF_FAMILIES];
  for (int f = 0; f < AZB_PROF_FAMILIES; ++f) {
    keep[f] = s->plan[f];
    pick[f] = s->plan[f];
    best[f] = 1e300;
  }
  int rc = AZB_OK;
  for (int p = 0; p < n && rc == AZB_OK; ++p) {
    int pl[AZB_PROF_FAMILIES];
    for (int f = 0; f < AZB_PROF_FAMILIES; ++f) pl[f] = (f == AZB_PROF_ROW || f == AZB_PROF_KSPACE) ? 0 : p;
    if ((rc = azb_set_plans(s, pl))) break;
    if ((rc = azb_dudt(s, u_hat, s->u1))) break;  // warm-up
    if ((rc = azb_profile_reset(s))) break;
    s->prof = true;
    for (int r = 0; r < reps && rc == AZB_OK; ++r) rc = azb_dudt(s, u_hat, s->u1);
    s->prof = false;
    if (rc) break;
    double ms[AZB_PROF_FAMILIES];
    long cnt[AZB_PROF_FAMILIES];
    if ((rc = azb_profile_read(s, ms, cnt, AZB_PROF_FAMILIES))) break;
    for (int f = 0; f < AZB_PROF_FAMILIES; ++f) {
      if (f == AZB_PROF_ROW || f == AZB_PROF_KSPACE || cnt[f] == 0) continue;
      const double avg = ms[f] / cnt[f];
      if (avg < best[f]) {
        best[f] = avg;
        pick[f] = p;
      }
    }
  }
  s->prof = prof_was;
  (void)azb_profile_reset(s);
  if (rc) {
    (void)azb_set_plans(s, keep);
    return rc;
  }
  return azb_set_plans(s, pick);
}

int azb_copy_bandwidth(double gib, int reps, double* gbs_out) {
  if (!gbs_out || !(gib > 0.0) || gib > 32.0 || reps < 1) return fail(AZB_ERR_ARG, "bad arguments");
  const long n = (long)(gib * 1073741824.0) / 16;
  void *x = nullptr, *y = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = AZB_OK;
  auto cleanup = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
  };
  if (hipMalloc(&x, n * 16) != hipSuccess || hipMalloc(&y, n * 16) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    cleanup();
    return fail(AZB_ERR_HIP, "allocation for the copy bandwidth probe failed");
  }
  const unsigned blocks = (unsigned)((n + 255) / 256);  // n < 2^31
  if (hipMemset(x, 0x3f, n * 16) != hipSuccess) {
    cleanup();
    return fail(AZB_ERR_HIP, "copy bandwidth probe: fill failed");
  }
  hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, nullptr, (const double2*)x, (double2*)y, n);
  (void)hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, nullptr, (const double2*)x, (double2*)y, n);
  (void)hipEventRecord(e1, nullptr);
  float ms = 0.f;
  if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f)
    rc = fail(AZB_ERR_HIP, "copy bandwidth probe failed");
  else
    *gbs_out = 2.0 * (double)n * 16.0 * reps / (ms * 1e-3) / 1e9;
  cleanup();
  return rc;
}

int azb_synchronize(azb_solver* s) {
  if (!s) return fail(AZB_ERR_ARG, "NULL solver");
  HIPCHK(hipStreamSynchronize(s->stream));
  return AZB_OK;
}

int azb_device_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return AZB_OK;
}

int azb_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(AZB_ERR_ARG, "NULL argument");
  HIPCHK(hipMalloc(ptr, bytes));
  return AZB_OK;
}
int azb_free(void* ptr) {
  HIPCHK(hipFree(ptr));
  return AZB_OK;
}
int azb_memcpy_h2d(void* dst, const void* src_host, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src_host, bytes, hipMemcpyHostToDevice));
  return AZB_OK;
}
int azb_memcpy_d2h(void* dst_host, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst_host, src, bytes, hipMemcpyDeviceToHost));
  return AZB_OK;
}
int azb_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return AZB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// standalone seams
// ---------------------------------------------------------------------------
namespace {
// work area of a standalone transform of length n: the device state the column
// kernels read (never marked done) at offset 0, the twiddle table after it
constexpr size_t kFftTwOffset = (sizeof(DevState) + 255) / 256 * 256;

template <typename T>
size_t fft_work_bytes(int n) {
  return kFftTwOffset + (sizeof(cx<T>) * (size_t)n + 255) / 256 * 256;
}

template <typename T>
int fft_work_init(void* work, int n) {
  const std::vector<cx<T>> h = host_twiddles<T>(n);
  std::vector<unsigned char> head(kFftTwOffset, 0);
  HIPCHK(hipMemcpy(work, head.data(), kFftTwOffset, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((char*)work + kFftTwOffset, h.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
  return AZB_OK;
}

// queue one rank-dim transform; no allocation, no synchronisation
template <typename T>
int fft_ws_T(int dim, int n, int howmany, bool forward, void* in, void* out, void* work, hipStream_t st) {
  if (!good_len(n)) return fail(AZB_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " not built");
  if (dim < 1 || dim > 3) return fail(AZB_ERR_ARG, "dim must be 1..3");
  const void* tw = (const char*)work + kFftTwOffset;
  const long nh = n / 2 + 1;
  long lines = howmany;
  for (int d = 1; d < dim; ++d) lines *= n;
  ColArgs c{};
  c.nsmall = n;
  c.tw = tw;
  c.st = (DevState*)work;
  PlainRowArgs p{};
  p.nlines = lines;
  p.tw = tw;
  int rc;
  // leading complex axes (axis d-2 .. 0), in place on the complex array
  auto leading = [&](void* arr, int dir) -> int {
    for (int ax = dim - 2; ax >= 0; --ax) {
      long inner = nh;  // product of extents after axis ax
      for (int d = ax + 1; d < dim - 1; ++d) inner *= n;
      long batches = howmany;
      for (int d = 0; d < ax; ++d) batches *= n;
      c.in = arr;
      c.out = arr;
      c.nq = 1;
      c.nk = (int)inner;
      c.in_qs = c.out_qs = 0;
      c.in_rs = c.out_rs = inner;
      c.in_bs = c.out_bs = inner * n;
      c.bps = (int)batches;  // a single device state: every batch maps to it
      int r = dir < 0 ? launch_col_full<T, -1>(n, c, batches, st) : launch_col_full<T, +1>(n, c, batches, st);
      if (r) return r;
    }
    return AZB_OK;
  };
  if (forward) {
    p.in = in;
    p.out = out;
    if ((rc = launch_plain_row<T, -1>(n, p, st))) return rc;
    if ((rc = leading(out, -1))) return rc;
  } else {
    if ((rc = leading(in, +1))) return rc;
    p.in = in;
    p.out = out;
    if ((rc = launch_plain_row<T, +1>(n, p, st))) return rc;
  }
  HIPCHK(hipGetLastError());
  return AZB_OK;
}

// the allocating form: a temporary work area, freed after the transform finished
template <typename T>
int fft_T(int dim, int n, int howmany, bool forward, void* in, void* out, hipStream_t st) {
  if (!good_len(n)) return fail(AZB_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " not built");
  if (dim < 1 || dim > 3) return fail(AZB_ERR_ARG, "dim must be 1..3");
  DevFree work;
  HIPCHK(hipMalloc(&work.p, fft_work_bytes<T>(n)));
  int rc = fft_work_init<T>(work.p, n);
  if (rc) return rc;
  if ((rc = fft_ws_T<T>(dim, n, howmany, forward, in, out, work.p, st))) return rc;
  HIPCHK(hipStreamSynchronize(st));
  return AZB_OK;
}

int fft_check(int precision, int howmany, const void* in, const void* out, int n) {
  if (!in || !out) return fail(AZB_ERR_ARG, "NULL argument");
  if (precision != AZB_FP64 && precision != AZB_FP32_MIXED) return fail(AZB_ERR_ARG, "unknown precision");
  if (howmany < 1) return fail(AZB_ERR_ARG, "howmany must be positive");
  if (!good_len(n)) return fail(AZB_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " not built");
  return AZB_OK;
}
}  // namespace

extern "C" {

int azb_fft_r2c(int dim, int n, int howmany, int precision, const void* in, void* out, void* stream) {
  int rc = fft_check(precision, howmany, in, out, n);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return precision == AZB_FP64 ? fft_T<double>(dim, n, howmany, true, (void*)in, out, st)
                               : fft_T<float>(dim, n, howmany, true, (void*)in, out, st);
}

int azb_fft_c2r(int dim, int n, int howmany, int precision, void* in, void* out, void* stream) {
  int rc = fft_check(precision, howmany, in, out, n);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return precision == AZB_FP64 ? fft_T<double>(dim, n, howmany, false, in, out, st)
                               : fft_T<float>(dim, n, howmany, false, in, out, st);
}

size_t azb_fft_work_bytes(int n, int precision) {
  if (!good_len(n) || (precision != AZB_FP64 && precision != AZB_FP32_MIXED)) return 0;
  return precision == AZB_FP64 ? fft_work_bytes<double>(n) : fft_work_bytes<float>(n);
}

int azb_fft_work_init(void* work, int n, int precision) {
  if (!work) return fail(AZB_ERR_ARG, "NULL work area");
  if (precision != AZB_FP64 && precision != AZB_FP32_MIXED) return fail(AZB_ERR_ARG, "unknown precision");
  if (!good_len(n)) return fail(AZB_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " not built");
  return precision == AZB_FP64 ? fft_work_init<double>(work, n) : fft_work_init<float>(work, n);
}

int azb_fft_r2c_ws(int dim, int n, int howmany, int precision, const void* in, void* out, void* work,
                   void* stream) {
  int rc = fft_check(precision, howmany, in, out, n);
  if (rc) return rc;
  if (!work) return fail(AZB_ERR_ARG, "NULL work area");
  hipStream_t st = (hipStream_t)stream;
  return precision == AZB_FP64 ? fft_ws_T<double>(dim, n, howmany, true, (void*)in, out, work, st)
                               : fft_ws_T<float>(dim, n, howmany, true, (void*)in, out, work, st);
}

int azb_fft_c2r_ws(int dim, int n, int howmany, int precision, void* in, void* out, void* work, void* stream) {
  int rc = fft_check(precision, howmany, in, out, n);
  if (rc) return rc;
  if (!work) return fail(AZB_ERR_ARG, "NULL work area");
  hipStream_t st = (hipStream_t)stream;
  return precision == AZB_FP64 ? fft_ws_T<double>(dim, n, howmany, false, in, out, work, st)
                               : fft_ws_T<float>(dim, n, howmany, false, in, out, work, st);
}

int azb_copy_to_padded(int dim, int N, int M, const void* in, void* out, void* stream) {
  if (!in || !out) return fail(AZB_ERR_ARG, "NULL argument");
  if ((dim != 2 && dim != 3) || M < N) return fail(AZB_ERR_ARG, "bad dim or sizes");
  const long mh = M / 2 + 1;
  const long total = dim == 2 ? (long)M * mh : (long)M * M * mh;
  hipLaunchKernelGGL(pad_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (const double2*)in, (double2*)out, dim, N, M);
  HIPCHK(hipGetLastError());
  return AZB_OK;
}

int azb_copy_from_padded(int dim, int M, int N, const void* in, void* out, void* stream) {
  if (!in || !out) return fail(AZB_ERR_ARG, "NULL argument");
  if ((dim != 2 && dim != 3) || N < 2 || M < N) return fail(AZB_ERR_ARG, "bad dim or sizes");
  const long nh = N / 2 + 1;
  const long total = dim == 2 ? (long)N * nh : (long)N * N * nh;
  hipLaunchKernelGGL(unpad_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (const double2*)in, (double2*)out, dim, M, N);
  HIPCHK(hipGetLastError());
  return AZB_OK;
}

int azb_scale(double a, void* x, long n, void* stream) {
  if (!x) return fail(AZB_ERR_ARG, "NULL argument");
  if (n < 0) return fail(AZB_ERR_ARG, "negative length");
  if (n == 0) return AZB_OK;
  hipLaunchKernelGGL(scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a,
                     (double*)x, n);
  HIPCHK(hipGetLastError());
  return AZB_OK;
}

int azb_leray(int dim, int N, void* u_hat, void* stream) {
  if (!u_hat) return fail(AZB_ERR_ARG, "NULL argument");
  if (dim != 2 && dim != 3) return fail(AZB_ERR_ARG, "dim must be 2 or 3");
  if (N < 2 || N % 2) return fail(AZB_ERR_ARG, "N must be even and positive");
  const long nh = N / 2 + 1;
  const long nmodes = dim == 2 ? (long)N * nh : (long)N * N * nh;
  hipLaunchKernelGGL(leray_kernel, dim3((unsigned)((nmodes + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (double2*)u_hat, dim, N, nmodes);
  HIPCHK(hipGetLastError());
  return AZB_OK;
}

int azb_energy_spectrum(int dim, int N, const void* u_hat, double* spec_host, void* stream) {
  if (!u_hat || !spec_host) return fail(AZB_ERR_ARG, "NULL argument");
  if (dim != 2 && dim != 3) return fail(AZB_ERR_ARG, "dim must be 2 or 3");
  if (N < 2 || N % 2) return fail(AZB_ERR_ARG, "N must be even and positive");
  hipStream_t st = (hipStream_t)stream;
  const long nh = N / 2 + 1;
  const long nmodes = dim == 2 ? (long)N * nh : (long)N * N * nh;
  DevFree buf;
  HIPCHK(hipMalloc(&buf.p, sizeof(double) * nh));
  double* d = (double*)buf.p;
  HIPCHK(hipMemsetAsync(d, 0, sizeof(double) * nh, st));
  long blocks = (nmodes + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(energy_spectrum_kernel, dim3((unsigned)blocks), dim3(256), sizeof(double) * nh, st,
                     (const double2*)u_hat, dim, N, nmodes, d);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(spec_host, d, sizeof(double) * nh, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return AZB_OK;
}

int azb_axpby(double a, const void* x, double b, void* y, long n, void* stream) {
  if (!x || !y) return fail(AZB_ERR_ARG, "NULL argument");
  if (n <= 0) return AZB_OK;
  hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a,
                     (const double2*)x, b, (double2*)y, n);
  HIPCHK(hipGetLastError());
  return AZB_OK;
}

int azb_max_norm(const void* x, long n, double* out_host, void* stream) {
  if (!x || !out_host) return fail(AZB_ERR_ARG, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  DevFree buf;
  HIPCHK(hipMalloc(&buf.p, sizeof(unsigned long long)));
  unsigned long long* d = (unsigned long long*)buf.p;
  HIPCHK(hipMemsetAsync(d, 0, sizeof(unsigned long long), st));
  if (n > 0) {
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(max_abs_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (const double*)x, n, d);
  }
  unsigned long long bits = 0;
  HIPCHK(hipMemcpyAsync(&bits, d, sizeof(bits), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double v;
  std::memcpy(&v, &bits, sizeof(v));
  *out_host = v;
  return AZB_OK;
}

}  // extern "C"
