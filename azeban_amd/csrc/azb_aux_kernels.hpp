// Non-template kernels of the solver's own translation unit: the per-step control
// kernels of the device-driven loop and the standalone seams (padding, unpadding,
// copy, scale, axpby, max-norm, Leray projection, energy spectrum).  Kept out of
// azb_kernels.hpp, which more than one translation unit includes.
#pragma once
#include "azb_kernels.hpp"

namespace azb {

// ---------------------------------------------------------------------------
// Small control kernels (single workgroup)
// ---------------------------------------------------------------------------
// clear the per-slot max bins; fixed-step mode sets max_dt from the host value
__global__ void step_prologue_fixed(DevState* st_all, double max_dt) {
  DevState* st = st_all + blockIdx.x;
  const int t = threadIdx.x;
  for (int e = t; e < 4 * kBins; e += blockDim.x) (&st->umax_bins[0][0])[e] = 0ull;
  if (t == 0) {
    st->done = 0;
    st->max_dt = max_dt;
  }
}

// device-driven loop: a step past t_end does nothing (done = 1)
__global__ void step_prologue_loop(DevState* st_all) {
  DevState* st = st_all + blockIdx.x;
  const int t = threadIdx.x;
  const bool go = st->time < st->t_end;
  for (int e = t; e < 4 * kBins; e += blockDim.x) (&st->umax_bins[0][0])[e] = 0ull;
  if (t == 0) {
    st->done = go ? 0 : 1;
    if (go) st->max_dt = fmin(st->t_end - st->time, st->max_dt_visc);
  }
}

__global__ void step_epilogue_loop(DevState* st_all) {
  DevState* st = st_all + blockIdx.x;
  if (threadIdx.x == 0 && !st->done) {
    st->time += st->dt;
    st->steps += 1;
  }
}

// (re)start the device-driven loop of one sample: its time, target and step count
__global__ void loop_set_sample(DevState* st_all, int sample, double time, double t_end, double max_dt_visc) {
  if (threadIdx.x != 0) return;
  DevState* st = st_all + sample;
  st->time = time;
  st->t_end = t_end;
  st->max_dt_visc = max_dt_visc;
  st->done = time < t_end ? 0 : 1;
  st->steps = 0;
}

// reduce one slot's bins: u_max = max / M^d; stage 1 also sets dt = min(C/(N u_max), max_dt)
__global__ void finalize_umax(DevState* st_all, int slot, int set_dt, double Md, double C, double N) {
  DevState* st = st_all + blockIdx.x;
  if (st->done) return;
  const int t = threadIdx.x;  // 64 threads
  double v = __longlong_as_double((long long)st->umax_bins[slot][t]);
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if (t == 0) {
    const double um = v / Md;
    st->u_max = um;
    if (set_dt) st->dt = fmin(C / (N * um), st->max_dt);
  }
}

// zero padding of one component (copy_to_padded_functions.hpp:11-63): full axes
// keep i <= N/2, move i >= N/2+1 to i + M - N, zero the gap; compact last axis
// keeps k < N/2+1.
__global__ void pad_kernel(const double2* in, double2* out, int dim, int N, int M) {
  const int nh = N / 2 + 1, mh = M / 2 + 1;
  const long total = dim == 2 ? (long)M * mh : (long)M * M * mh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int k = (int)(idx % mh);
  long rest = idx / mh;
  auto src_full = [&](int i) -> int {
    if (i < N / 2 + 1) return i;
    if (i < N / 2 + 1 + M - N) return -1;
    return i + N - M;
  };
  double2 r = make_double2(0.0, 0.0);
  if (dim == 2) {
    const int i = src_full((int)rest);
    if (i >= 0 && k < nh) r = in[(long)i * nh + k];
  } else {
    const int j = src_full((int)(rest % M));
    const int i = src_full((int)(rest / M));
    if (i >= 0 && j >= 0 && k < nh) r = in[((long)i * N + j) * nh + k];
  }
  out[idx] = r;
}

// truncation of one component M -> N (copy_from_padded_functions.hpp:11-27,
// copy_from_padded.cpp:28-86): full axes read i < N/2+1 from i and the rest from
// i + M - N; the compact last axis reads k from k.
__global__ void unpad_kernel(const double2* in, double2* out, int dim, int M, int N) {
  const int nh = N / 2 + 1, mh = M / 2 + 1;
  const long total = dim == 2 ? (long)N * nh : (long)N * N * nh;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int k = (int)(idx % nh);
  const long rest = idx / nh;
  auto src_full = [&](int i) -> long { return i < N / 2 + 1 ? i : i + M - N; };
  if (dim == 2) {
    out[idx] = in[src_full((int)rest) * mh + k];
  } else {
    const long j = src_full((int)(rest % N));
    const long i = src_full((int)(rest / N));
    out[idx] = in[(i * M + j) * mh + k];
  }
}

// Streaming copy for the bandwidth reference over n 16-byte elements.
__global__ __launch_bounds__(256) void stream_copy_kernel(const double2* __restrict__ x, double2* __restrict__ y,
                                                          long n) {
  // one 16-byte element per thread, consecutive workgroups on consecutive 4 KiB:
  // on MI355X this reaches ~6.3 TB/s where a grid-stride loop with four loads in
  // flight per thread stays near 4.7 TB/s
  using v2 = vec2_t<double>;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(((const v2*)x)[i], (v2*)y + i);
}

// x <- a x over a flat real array (scale.hpp)
__global__ void scale_kernel(double a, double* x, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}

// y <- a x + b y over a flat complex array (axpby.hpp:41-46)
__global__ void axpby_kernel(double a, const double2* x, double b, double2* y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 xv = x[i], yv = y[i];
  y[i] = make_double2(a * xv.x + b * yv.x, a * xv.y + b * yv.y);
}

// max |x| over a flat real array (norm.hpp:72-91) into a device scalar
__global__ void max_abs_kernel(const double* x, long n, unsigned long long* out) {
  double m = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    m = fmax(m, fabs(x[i]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

}  // namespace azb

namespace azb {

// Leray projection in place, u_hat <- (I - k k^T / |k|^2) u_hat, k = 0 mode set to 0
// (src/azeban/operations/leray.cpp:28-116).
__global__ void leray_kernel(double2* u, int dim, int N, long nmodes) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nmodes) return;
  const int nh = N / 2 + 1;
  const double twopi = 2.0 * 3.14159265358979323846;
  const int kk = (int)(idx % nh);
  const long rest = idx / nh;
  double k[3];
  if (dim == 2) {
    const int i = (int)rest;
    k[0] = twopi * (i >= nh ? i - N : i);
    k[1] = twopi * kk;
    k[2] = 0;
  } else {
    const int j = (int)(rest % N), i = (int)(rest / N);
    k[0] = twopi * (i >= nh ? i - N : i);
    k[1] = twopi * (j >= nh ? j - N : j);
    k[2] = twopi * kk;
  }
  double absk2 = 0;
  for (int d = 0; d < dim; ++d) absk2 += k[d] * k[d];
  double2 v[3];
  for (int d = 0; d < dim; ++d) v[d] = u[(long)d * nmodes + idx];
  for (int r = 0; r < dim; ++r) {
    double2 o = make_double2(0.0, 0.0);
    if (absk2 != 0) {
      for (int c = 0; c < dim; ++c) {
        const double p = (r == c ? 1.0 : 0.0) - (k[r] * k[c]) / absk2;
        o.x += p * v[c].x;
        o.y += p * v[c].y;
      }
    }
    u[(long)r * nmodes + idx] = o;
  }
}

// Energy spectrum E(K), K = max|k_i| (L-infinity shells), 1/2 |u_hat / N^d|^2 with
// conjugate doubling for k_last != 0 (src/azeban/operations/energy_spectrum.cpp:10-50,
// include/azeban/operations/spectrum.hpp:33-118).  Shell sums accumulate in LDS per
// workgroup, then one double atomicAdd per shell per workgroup.
__global__ void energy_spectrum_kernel(const double2* u, int dim, int N, long nmodes, double* spec) {
  extern __shared__ double sh[];
  const int nh = N / 2 + 1;
  for (int e = threadIdx.x; e < nh; e += blockDim.x) sh[e] = 0.0;
  __syncthreads();
  const double norm = 1.0 / pow((double)N, dim);
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < nmodes;
       idx += (long)gridDim.x * blockDim.x) {
    const int kk = (int)(idx % nh);
    const long rest = idx / nh;
    int K;
    if (dim == 2) {
      const int i = (int)rest;
      const int k1 = abs(i >= nh ? i - N : i);
      K = max(k1, kk);
    } else {
      const int j = (int)(rest % N), i = (int)(rest / N);
      const int k1 = abs(i >= nh ? i - N : i), k2 = abs(j >= nh ? j - N : j);
      K = max(max(k1, k2), kk);
    }
    double e = 0;
    for (int d = 0; d < dim; ++d) {
      const double2 v = u[(long)d * nmodes + idx];
      const double x = norm * v.x, y = norm * v.y;
      e += x * x + y * y;
    }
    e *= 0.5;
    if (kk != 0) e *= 2.0;
    atomicAdd(&sh[K], e);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nh; e += blockDim.x)
    if (sh[e] != 0.0) atomicAdd(&spec[e], sh[e]);
}

}  // namespace azb
