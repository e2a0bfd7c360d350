// Device-side FFT building blocks for gfx950: complex type, small-radix DFTs in
// registers, and a workgroup-wide Stockham FFT whose inter-pass exchange goes
// through LDS.
//
// A "block FFT" transforms TS independent length-L sequences held by NT threads
// (E = TS*L/NT complex values per thread).  Before the pass of radix R, thread t
// holds, for every butterfly b in [0, E/R):
//     g = t + NT*b,  s = g % TS (sequence),  j = g / TS (butterfly),
//     v[b*R + r] = x_s[j + r*L/R],  r in [0, R).
// After the last pass the same formula (with the last radix) names the natural-
// order outputs each thread holds, so callers load the first pass straight from
// HBM and store the last pass straight to HBM without an LDS round trip.
//
// Stockham autosort step (pass with accumulated size Ns, radix R):
//     k = j mod Ns,  v_r *= W_{Ns R}^{r k},  V = DFT_R(v),
//     y[(j - k) R + k + r Ns] = V_r.
// Twiddles come from a per-length table tw[t] = exp(-2 pi i t / L).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace azb {

template <typename T>
struct cx {
  T x, y;
};

// Complex elements as a native two-vector, for 8/16-byte loads and stores through
// pointers in the global address space: a uniform (scalar) base plus a 32-bit
// element offset, with no 64-bit address arithmetic per element.
template <typename T>
using vec2_t = T __attribute__((ext_vector_type(2)));
template <typename T>
using gvec2 = __attribute__((address_space(1))) vec2_t<T>;

template <typename T>
__device__ __forceinline__ cx<T> gload(const void* base, unsigned idx) {
  const vec2_t<T> v = ((const gvec2<T>*)base)[idx];
  cx<T> r;
  r.x = v.x;
  r.y = v.y;
  return r;
}
// Stores are non-temporal: a column pass never reads its own output, and with
// non-temporal stores the column passes measured 5-9 % faster.  Loads stay cached
// (neighbouring tiles share the lines that straddle tile boundaries; non-temporal
// loads measured 7-25 % slower).
template <typename T>
__device__ __forceinline__ void gstore(void* base, unsigned idx, T x, T y) {
  vec2_t<T> v;
  v.x = x;
  v.y = y;
  __builtin_nontemporal_store(v, &((gvec2<T>*)base)[idx]);
}

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) {
  cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) {
  return mk(a.x + b.x, a.y + b.y);
}
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) {
  return mk(a.x - b.x, a.y - b.y);
}
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a b with the two fused multiply-adds fixed in the source, so that every
// instantiation (any tile width or thread mapping) rounds identically
template <typename T>
__device__ __forceinline__ cx<T> cmul_fixed(cx<T> a, cx<T> b) {
  const T yy = a.y * b.y, yx = a.y * b.x;
  return mk(T(fma(a.x, b.x, -yy)), T(fma(a.x, b.y, yx)));
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) {
  return mk(a.x * s, a.y * s);
}
// multiply by i*sgn (sgn = +1 or -1)
template <int SGN, typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) {
  return SGN > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// constexpr cos/sin of 2*pi*m/n for compile-time twiddles inside composite radices
// ---------------------------------------------------------------------------
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double ce_sin_poly(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2 * n) * (2 * n + 1));
    sum += term;
  }
  return sum;
}
constexpr double ce_cos_poly(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}
// returns cos(2 pi m / n) (q = 0) or sin(2 pi m / n) (q = 1) via octant reduction
constexpr double ce_trig(long m, long n, int q) {
  m %= n;
  if (m < 0) m += n;
  // angle = 2 pi m / n ; work in units of n/8 octants using exact integer math
  long m8 = 8 * m;
  int oct = (int)(m8 / n);  // 0..7
  double rem = (double)(m8 - (long)oct * n) / (double)n;  // [0,1)
  double phi = rem * (kPi / 4);  // within octant
  double c = 0, s = 0;
  // angle = oct*pi/4 + phi
  double cphi = ce_cos_poly(phi), sphi = ce_sin_poly(phi);
  const double h = 0.70710678118654752440084436210484903928;
  switch (oct) {
    case 0: c = cphi; s = sphi; break;
    case 1: c = h * (cphi - sphi); s = h * (cphi + sphi); break;
    case 2: c = -sphi; s = cphi; break;
    case 3: c = -h * (cphi + sphi); s = h * (cphi - sphi); break;
    case 4: c = -cphi; s = -sphi; break;
    case 5: c = -h * (cphi - sphi); s = -h * (cphi + sphi); break;
    case 6: c = sphi; s = -cphi; break;
    default: c = h * (cphi + sphi); s = -h * (cphi - sphi); break;
  }
  return q == 0 ? c : s;
}

// ---------------------------------------------------------------------------
// In-register DFTs, natural-order in and out.  DIR = -1 forward, +1 inverse.
// ---------------------------------------------------------------------------
template <int R, int DIR, typename T>
struct DFT;

template <int DIR, typename T>
struct DFT<1, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>*) {}
};

template <int DIR, typename T>
struct DFT<2, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <int DIR, typename T>
struct DFT<3, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    // c = DIR sin(2 pi / 3); v1, v2 = t2 +- i c d with the products fused into the
    // sums (two operations fewer than forming i c d first)
    const T c = T(DIR) * T(0.86602540378443864676372317075293618);
    cx<T> t1 = v[1] + v[2];
    cx<T> t2 = mk(v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y);
    cx<T> d = v[1] - v[2];
    v[0] = v[0] + t1;
    v[1] = mk(fma(-c, d.y, t2.x), fma(c, d.x, t2.y));
    v[2] = mk(fma(c, d.y, t2.x), fma(-c, d.x, t2.y));
  }
};

template <int DIR, typename T>
struct DFT<4, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> a = v[0] + v[2], b = v[0] - v[2];
    cx<T> c = v[1] + v[3], d = mul_i<DIR>(v[1] - v[3]);
    v[0] = a + c;
    v[2] = a - c;
    v[1] = b + d;
    v[3] = b - d;
  }
};

template <int DIR, typename T>
struct DFT<8, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>* v) {
    const T h = T(0.70710678118654752440084436210484903928);
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    DFT<4, DIR, T>::run(e);
    DFT<4, DIR, T>::run(o);
    // W8^1 = h(1 + DIR i), W8^2 = DIR i, W8^3 = h(-1 + DIR i); the factor h is
    // fused into the output sums
    cx<T> s1 = o[1] + mul_i<DIR>(o[1]);
    cx<T> o2 = mul_i<DIR>(o[2]);
    cx<T> s3 = mul_i<DIR>(o[3]) - o[3];
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = mk(fma(h, s1.x, e[1].x), fma(h, s1.y, e[1].y));
    v[5] = mk(fma(-h, s1.x, e[1].x), fma(-h, s1.y, e[1].y));
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = mk(fma(h, s3.x, e[3].x), fma(h, s3.y, e[3].y));
    v[7] = mk(fma(-h, s3.x, e[3].x), fma(-h, s3.y, e[3].y));
  }
};

// compile-time table of exp(DIR 2 pi i e / R), e in [0, R)
template <int R, int DIR>
struct RootTable {
  double c[R];
  double s[R];
  constexpr RootTable() : c(), s() {
    for (int e = 0; e < R; ++e) {
      c[e] = ce_trig(e, R, 0);
      s[e] = DIR * ce_trig(e, R, 1);
    }
  }
};

// Composite radix R = P*Q: n = Q p + q, k = k1 + P k2
//   X[k1 + P k2] = sum_q W_Q^{q k2} W_R^{q k1} sum_p x[Q p + q] W_P^{p k1}
template <int P, int Q, int DIR, typename T>
struct DFTComposite {
  __device__ __forceinline__ static void run(cx<T>* v) {
    constexpr int R = P * Q;
    constexpr RootTable<R, DIR> tab{};
    cx<T> y[Q][P];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
#pragma unroll
      for (int p = 0; p < P; ++p) y[q][p] = v[Q * p + q];
      DFT<P, DIR, T>::run(y[q]);
#pragma unroll
      for (int k1 = 1; k1 < P; ++k1) {
        if (q != 0) y[q][k1] = cmul(y[q][k1], mk(T(tab.c[(q * k1) % R]), T(tab.s[(q * k1) % R])));
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      cx<T> z[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) z[q] = y[q][k1];
      DFT<Q, DIR, T>::run(z);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = z[k2];
    }
  }
};

// Prime-factor (Good-Thomas) radix R = P*Q with gcd(P, Q) = 1: input index
// n = (Q n1 + P n2) mod R, output index k = (A k1 + B k2) mod R with A = 1 (mod P),
// 0 (mod Q) and B = 0 (mod P), 1 (mod Q).  Then n k = Q n1 k1 + P n2 k2 (mod R), so
//   X[A k1 + B k2] = sum_n2 W_Q^{n2 k2} sum_n1 W_P^{n1 k1} x[Q n1 + P n2]
// needs no twiddles between the two DFT stages (the index maps are register renames).
constexpr int ce_inverse_mod(int a, int m) {
  for (int x = 1; x <= m; ++x)
    if ((a * x) % m == 1) return x;
  return 0;
}
template <int P, int Q, int DIR, typename T>
struct DFTPrimeFactor {
  static constexpr int R = P * Q;
  static constexpr int A = (Q * ce_inverse_mod(Q % P, P)) % R;
  static constexpr int B = (P * ce_inverse_mod(P % Q, Q)) % R;
  static_assert(A % P == 1 % P && A % Q == 0 && B % P == 0 && B % Q == 1 % Q, "P and Q must be coprime");
  __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> y[Q][P];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) y[n2][n1] = v[(Q * n1 + P * n2) % R];
      DFT<P, DIR, T>::run(y[n2]);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      cx<T> z[Q];
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) z[n2] = y[n2][k1];
      DFT<Q, DIR, T>::run(z);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[(A * k1 + B * k2) % R] = z[k2];
    }
  }
};

template <int DIR, typename T>
struct DFT<6, DIR, T> : DFTPrimeFactor<2, 3, DIR, T> {};
template <int DIR, typename T>
struct DFT<12, DIR, T> : DFTPrimeFactor<4, 3, DIR, T> {};
template <int DIR, typename T>
struct DFT<16, DIR, T> : DFTComposite<4, 4, DIR, T> {};
template <int DIR, typename T>
struct DFT<24, DIR, T> : DFTPrimeFactor<8, 3, DIR, T> {};

// ---------------------------------------------------------------------------
// Workgroup Stockham FFT
// ---------------------------------------------------------------------------
template <int... Rs>
struct Plan {};

template <int R0, int... Rs>
struct PlanFirst {
  static constexpr int value = R0;
};
template <int R0>
constexpr int plan_last() { return R0; }
template <int R0, int R1, int... Rs>
constexpr int plan_last() { return plan_last<R1, Rs...>(); }
template <int R0>
constexpr long plan_prod() { return R0; }
template <int R0, int R1, int... Rs>
constexpr long plan_prod() { return (long)R0 * plan_prod<R1, Rs...>(); }

// LDS words needed by the exchange (SPLIT: one T per element, else two)
template <typename T, int L, int TS, bool SPLIT>
struct ExchangeSize {
  static constexpr int words = TS * L * (SPLIT ? 1 : 2);
};

template <typename T, int L, int TS, int NT, int DIR, bool SPLIT>
struct BlockFFT {
  static constexpr int E = TS * L / NT;
  static_assert(TS * L % NT == 0, "TS*L must be a multiple of NT");

  // twiddle W_L^{e} with direction applied; the table pointer is in the global
  // address space: global (not flat) loads
  // with a scalar base and a 32-bit offset, and no flat-load entries in the LDS
  // wait counter
  using twptr = const __attribute__((address_space(1))) cx<T>*;

  __device__ __forceinline__ static cx<T> twiddle(twptr tw, unsigned e) {
    cx<T> w;
    w.x = tw[e].x;
    w.y = DIR > 0 ? -tw[e].y : tw[e].y;
    return w;
  }

  template <int Ns, int R>
  __device__ __forceinline__ static void butterflies(cx<T>* v, int t, const cx<T>* __restrict__ tw) {
    static_assert(E % R == 0, "radix must divide E");
#pragma unroll
    for (int b = 0; b < E / R; ++b) {
      cx<T>* w = v + b * R;
      if (Ns > 1) {
        const unsigned g = (unsigned)t + NT * b;
        const unsigned j = g / TS;
        const unsigned k = j % Ns;
        constexpr unsigned step = L / (Ns * R);
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = cmul(w[r], twiddle((twptr)tw, r * k * step));
      }
      DFT<R, DIR, T>::run(w);
    }
  }

  // write pass (Ns,R) outputs into LDS in [pos][s] order, then read the next pass's inputs
  template <int Ns, int R, int Rn>
  __device__ __forceinline__ static void exchange(cx<T>* v, int t, T* lds, bool active) {
    constexpr int comps = SPLIT ? 2 : 1;
#pragma unroll
    for (int c = 0; c < comps; ++c) {
      __syncthreads();
      if (active) {
#pragma unroll
        for (int b = 0; b < E / R; ++b) {
          const unsigned g = (unsigned)t + NT * b;
          const unsigned s = g % TS;
          const unsigned j = g / TS;
          const unsigned k = j % Ns;
          const unsigned base = (j - k) * R + k;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const unsigned idx = (base + r * Ns) * TS + s;
            if (SPLIT) {
              lds[idx] = c == 0 ? v[b * R + r].x : v[b * R + r].y;
            } else {
              lds[2 * idx] = v[b * R + r].x;
              lds[2 * idx + 1] = v[b * R + r].y;
            }
          }
        }
      }
      __syncthreads();
      if (active) {
#pragma unroll
        for (int b = 0; b < E / Rn; ++b) {
          const unsigned g = (unsigned)t + NT * b;
          const unsigned s = g % TS;
          const unsigned j = g / TS;
#pragma unroll
          for (int r = 0; r < Rn; ++r) {
            const unsigned idx = (j + r * (L / Rn)) * TS + s;
            if (SPLIT) {
              if (c == 0)
                v[b * Rn + r].x = lds[idx];
              else
                v[b * Rn + r].y = lds[idx];
            } else {
              v[b * Rn + r].x = lds[2 * idx];
              v[b * Rn + r].y = lds[2 * idx + 1];
            }
          }
        }
      }
    }
  }

  template <int Ns, int R>
  __device__ __forceinline__ static void passes(cx<T>* v, int t, T* lds, const cx<T>* tw,
                                                bool active, Plan<R>) {
    if (active) butterflies<Ns, R>(v, t, tw);
  }
  template <int Ns, int R, int Rn, int... Rs>
  __device__ __forceinline__ static void passes(cx<T>* v, int t, T* lds, const cx<T>* tw,
                                                bool active, Plan<R, Rn, Rs...>) {
    if (active) butterflies<Ns, R>(v, t, tw);
    exchange<Ns, R, Rn>(v, t, lds, active);
    passes<Ns * R, Rn, Rs...>(v, t, lds, tw, active, Plan<Rn, Rs...>{});
  }

  // Run the whole transform.  All threads of the workgroup must call this (it holds
  // barriers); threads with active == false only take part in the barriers.
  template <int... Rs>
  __device__ __forceinline__ static void run(cx<T>* v, int t, T* lds, const cx<T>* tw, bool active,
                                             Plan<Rs...> p) {
    static_assert(plan_prod<Rs...>() == L, "plan must multiply to L");
    passes<1, Rs...>(v, t, lds, tw, active, p);
  }

  // first-pass load: ld(s, pos) -> cx<T>
  template <int R0, class Ld>
  __device__ __forceinline__ static void load(cx<T>* v, int t, Ld ld) {
#pragma unroll
    for (int b = 0; b < E / R0; ++b) {
      const unsigned g = (unsigned)t + NT * b;
      const unsigned s = g % TS;
      const unsigned j = g / TS;
#pragma unroll
      for (int r = 0; r < R0; ++r) v[b * R0 + r] = ld(s, j + r * (L / R0));
    }
  }
  // last-pass store: st(s, pos, value)
  template <int RL, class St>
  __device__ __forceinline__ static void store(const cx<T>* v, int t, St st) {
#pragma unroll
    for (int b = 0; b < E / RL; ++b) {
      const unsigned g = (unsigned)t + NT * b;
      const unsigned s = g % TS;
      const unsigned j = g / TS;
#pragma unroll
      for (int r = 0; r < RL; ++r) st(s, j + r * (L / RL), v[b * RL + r]);
    }
  }
};

}  // namespace azb

namespace azb {

// ---------------------------------------------------------------------------
// Workgroup Stockham FFT with a per-pass thread <-> butterfly mapping.
//   MAP_STD : g = t + NT b, s = g % TS, j = g / TS (sequences fastest across threads)
//   MAP_SEQ : a thread owns one butterfly j of G consecutive sequences at once, so
//             after the last pass it holds the same positions of all G sequences
//             (elementwise products across sequences need no exchange), and a first
//             pass reading HBM is contiguous in j across threads
//   MAP_SEQ1: MAP_SEQ with G = 1 whatever the transform's G: consecutive threads own
//             consecutive butterflies of one sequence (a radix-L/64 pass puts each
//             sequence in exactly one wave, butterfly j in lane j)
//   MAP_WAVE: wave w owns sequence w (NT = 64 TS and L = 64 E), lane l its
//             butterflies l + 64 b
//   MAP_HALF: half-wave h (32 consecutive threads) owns sequence h (NT = 32 TS and
//             L = 32 E), lane l of it the butterflies l + 32 b, so wave w holds
//             sequences 2w and 2w + 1
// An exchange between two passes that keep every sequence inside the same wave
// (MAP_WAVE, or MAP_SEQ1 with L/R = 64: wave w holds sequence w; MAP_HALF, or
// MAP_SEQ1 with L/R = 32: wave w holds sequences 2w, 2w + 1) moves data only
// within a wave and needs no workgroup barrier.  Any mapping can be used by any
// pass: the LDS exchange writes with the mapping of the pass that produced the
// data and reads with the mapping of the next pass.
// ---------------------------------------------------------------------------
enum { MAP_STD = 0, MAP_SEQ = 1, MAP_SEQ1 = 3, MAP_WAVE = 5, MAP_HALF = 6 };

// Index arithmetic is unsigned throughout: signed division and modulo by a power of
// two cost a sign fix-up each, and these run once per element of every exchange.
template <int MAP, int TS, int NT, int L, int R, int G>
__device__ __forceinline__ void bmap(unsigned t, unsigned b, unsigned& s, unsigned& j) {
  if constexpr (MAP == MAP_STD) {
    if constexpr (NT % TS == 0) {  // the same sequence for every butterfly of a thread
      s = t % TS;
      j = t / TS + (unsigned)(NT / TS) * b;
    } else {
      const unsigned g = t + NT * b;
      s = g % TS;
      j = g / TS;
    }
  } else if constexpr (MAP == MAP_WAVE) {
    static_assert(NT % 64 == 0, "whole waves");
    s = t / 64u;
    j = t % 64u + 64u * b;
  } else if constexpr (MAP == MAP_HALF) {
    static_assert(NT == 32 * TS, "one half-wave per sequence");
    s = t / 32u;
    j = t % 32u + 32u * b;
  } else if constexpr (MAP == MAP_SEQ || MAP == MAP_SEQ1) {
    constexpr unsigned GG = MAP == MAP_SEQ1 ? 1 : G;
    const unsigned q = b / GG, gg = b % GG;
    const unsigned u = t + NT * q;
    j = u % (unsigned)(L / R);
    s = (u / (unsigned)(L / R)) * GG + gg;
  } else {
    static_assert(MAP != MAP, "unknown mapping");
  }
}

template <int... Rs>
struct PlanAt;
template <int R0, int... Rs>
struct PlanAt<R0, Rs...> {
  static constexpr int count = 1 + sizeof...(Rs);
};

// LDS layout of the exchange.  SWZ = false: element (s, pos) at pos*TS + s.
// SWZ = true: sequence-major with one pad slot per 256-byte bank row and a
// sequence stride of SOFF modulo the row, so lane strides of 1, TS or R*TS in
// either index stay spread over the banks (SOFF = 1 suits six interleaved
// sequences; three interleaved sequences with consecutive positions want the
// three runs of five or six slots apart, SOFF = 11).
template <typename T, int L, int TS, bool SWZ, int SOFF = 1, int PADE_ = 0>
struct ExLayout {
  // elements per 256-byte bank row (a pad slot follows each); a split exchange
  // (real and imaginary parts in turn) stores one T per slot
  static constexpr int PADE = PADE_ ? PADE_ : 256 / (2 * (int)sizeof(T));
  static constexpr int LP = ((L + L / PADE + PADE - 1) / PADE) * PADE + SOFF;
  static constexpr int elems = SWZ ? TS * LP : TS * L;
  __device__ __forceinline__ static unsigned at(unsigned s, unsigned pos) {
    return SWZ ? s * (unsigned)LP + pos + pos / (unsigned)PADE : pos * (unsigned)TS + s;
  }
  // at(s, pos + D) - at(s, pos) when that does not depend on pos (D a multiple of PADE)
  static constexpr bool linear(int D) { return !SWZ || D % PADE == 0; }
  static constexpr unsigned delta(int D) { return SWZ ? D + D / PADE : D * TS; }
};

// Sequence-major layout with an XOR swizzle and no pad slots: element (s, pos) at
// s*LP + (pos ^ (((pos >> A) & MASK) << S)).  The swizzle permutes positions inside
// aligned blocks, so it is a bijection on [0, L) that stays inside the sequence's
// LP slots; a position step that is a multiple of (MASK+1) << A leaves the swizzle
// term unchanged, so such steps are constant address offsets.
template <typename T, int L, int TS, int A, int MASK, int S, int LP_>
struct XorLayout {
  // S < A: each swizzled bit takes its term from a higher bit, so the map can be
  // undone from the top bit down (a bijection); S = A would clear bits instead
  static_assert((MASK & (MASK + 1)) == 0 && S < A && L % ((MASK + 1) << A) == 0 && LP_ >= L,
                "swizzle blocks must tile a sequence");
  static constexpr int LP = LP_;
  static constexpr int elems = TS * LP;
  __device__ __forceinline__ static unsigned at(unsigned s, unsigned pos) {
    return s * (unsigned)LP + (pos ^ (((pos >> A) & (unsigned)MASK) << S));
  }
  static constexpr bool linear(int D) { return D % ((MASK + 1) << A) == 0; }
  static constexpr unsigned delta(int D) { return D; }
};

// Layout of exchange P (the one after pass P): LSEL::lay<P> when a selector is
// given, else the transform's default layout D.
template <class LSEL, int P, class D>
struct LaySel {
  using type = typename LSEL::template lay<P>;
};
template <int P, class D>
struct LaySel<void, P, D> {
  using type = D;
};
template <class LSEL, class D>
struct LayElems {
  static constexpr int value = LSEL::elems > D::elems ? LSEL::elems : D::elems;
};
template <class D>
struct LayElems<void, D> {
  static constexpr int value = D::elems;
};

// MIDW / MIDH: bit P set = middle pass P uses MAP_WAVE / MAP_HALF instead of
// MAP_STD.  PADE: pad period of the swizzled layout in elements (0: one 16-byte
// complex per slot).  LSEL: per-exchange layouts (void: ExLayout for every
// exchange); every layout of one transform must give each sequence the same LP
// slots, so that a wave-local exchange writes only where the previous exchange's
// readers of its own sequences read.
// PTAB: tw is this transform's per-pass twiddle table (pass_table_entries /
// fill_pass_table: for each pass P >= 1 and each of its Ns twiddle classes k one
// row of R entries, W^{r k L/(Ns R)} with the transform's direction applied, for
// r = 1 .. R-1 at row[r-1]) instead of the W_L^e table: one base address per
// butterfly, constant offsets, 16-byte loads of two FP32 entries.
template <typename T, int L, int TS, int NT, int DIR, bool SPLIT, int G, int MAPF, int MAPL, bool SWZ = false,
          int SOFF = 1, unsigned MIDW = 0, int PADE = 0, unsigned MIDH = 0, class LSEL = void, bool PTAB = false>
struct BlockFFT2 {
  static constexpr int E = TS * L / NT;
  static_assert(TS * L % NT == 0, "TS*L must be a multiple of NT");
  static_assert(MIDW == 0 || (NT == 64 * TS && L == 64 * E), "MAP_WAVE needs one wave per sequence");
  static_assert(MIDH == 0 || (NT == 32 * TS && L == 32 * E), "MAP_HALF needs one half-wave per sequence");
  static_assert((MIDW & MIDH) == 0, "one mapping per pass");
  using LY = ExLayout<T, L, TS, SWZ, SOFF, PADE>;
  template <int P>
  using XL = typename LaySel<LSEL, P, LY>::type;
  // LDS words (of T) the exchange needs
  static constexpr int lds_words = SPLIT ? LayElems<LSEL, LY>::value : 2 * LayElems<LSEL, LY>::value;

  template <int P, int NP>
  static constexpr int map_of() {
    return P == 0 ? MAPF
                  : (P == NP - 1 ? MAPL
                                 : (((MIDW >> P) & 1u) ? MAP_WAVE : (((MIDH >> P) & 1u) ? MAP_HALF : MAP_STD)));
  }
  // sequences per wave when this mapping keeps every sequence inside one wave
  // (wave w holds sequences spw w ... spw w + spw - 1), else 0
  template <int MAP, int R>
  static constexpr int wave_owner() {
    if (MAP == MAP_WAVE || (MAP == MAP_SEQ1 && L / R == 64 && E == R)) return NT == 64 * TS ? 1 : 0;
    if (MAP == MAP_HALF || (MAP == MAP_SEQ1 && L / R == 32 && E == R)) return NT == 32 * TS ? 2 : 0;
    return 0;
  }
  // Between the write and the read phases of an exchange, and before the writes
  // (earlier reads of the same words): a workgroup barrier, or, when both mappings
  // keep each sequence in its own wave, only this wave's LDS operations completing
  // (a wave's LDS instructions execute in order; the wait also keeps the compiler
  // from moving LDS accesses across it).  A wave-local exchange follows a pass whose
  // mapping is wave-local, so the previous exchange's reads were this wave's own.
  template <bool LOCAL>
  __device__ __forceinline__ static void exchange_sync() {
    if constexpr (LOCAL)
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    else
      __syncthreads();
  }

  // twiddle table pointer in the global address space: global (not flat) loads
  // with a scalar base and a 32-bit offset, and no flat-load entries in the LDS
  // wait counter
  using twptr = const __attribute__((address_space(1))) cx<T>*;

  __device__ __forceinline__ static cx<T> twiddle(twptr tw, unsigned e) {
    cx<T> w;
    w.x = tw[e].x;
    w.y = DIR > 0 ? -tw[e].y : tw[e].y;
    return w;
  }

  template <int MAP, int Ns, int R>
  __device__ __forceinline__ static void butterflies(cx<T>* v, int t_in, const cx<T>* __restrict__ tw_in) {
    static_assert(E % R == 0, "radix must divide E");
    // keep this pass's twiddle loads inside the pass: without the opaque copy the
    // compiler hoists the (data-independent) loads of every pass of every transform
    // in the kernel to its top and holds them all in registers
    twptr tw = (twptr)tw_in;
    asm volatile("" : "+s"(tw));
    unsigned t = t_in;
    asm volatile("" : "+v"(t));
    __builtin_assume(t < (unsigned)NT);
#pragma unroll
    for (int b = 0; b < E / R; ++b) {
      cx<T>* w = v + b * R;
      if (Ns > 1) {
        unsigned s, j;
        bmap<MAP, TS, NT, L, R, G>(t, b, s, j);
        const unsigned k = j % Ns;
        constexpr unsigned step = L / (Ns * R);
        if constexpr (PTAB) {
          static_assert(sizeof(T) == 4 && R % 2 == 0, "per-pass tables: FP32, even radix");
          // row k of this pass's table, 16-byte aligned (rows of R entries); entries
          // r = 1 .. R-1 at row[r-1], loaded two at a time
          typedef float f4 __attribute__((ext_vector_type(4)));
          const __attribute__((address_space(1))) f4* row =
              (const __attribute__((address_space(1))) f4*)(tw + k * (unsigned)R);
#pragma unroll
          for (int q = 0; q < R / 2; ++q) {
            const f4 two = row[q];
            w[2 * q + 1] = cmul(w[2 * q + 1], mk(T(two.x), T(two.y)));
            if (2 * q + 2 < R) w[2 * q + 2] = cmul(w[2 * q + 2], mk(T(two.z), T(two.w)));
          }
        } else if (sizeof(T) == 8) {
          // FP64: one table load per butterfly, the other powers by repeated
          // multiplication (w^(r+1) = w^r w, below 22 ulp at R = 24).  Each saved
          // load is a full vector-memory instruction through the texture path; the
          // chain costs four FMA-class operations and keeps two powers live.
          // Loading every power instead made the 768-point FP64 row kernel 12 %
          // slower (6.39 against 5.72 ms per 512^3 launch; column passes unchanged).
          const cx<T> w1 = twiddle(tw, k * step);
          cx<T> wr = w1;
          w[1] = cmul_fixed(w[1], w1);
#pragma unroll
          for (int r = 2; r < R; ++r) {
            wr = cmul_fixed(wr, w1);
            w[r] = cmul_fixed(w[r], wr);
          }
        } else {
          // FP32 loads every power: a chain's error would rival the transform's own
#pragma unroll
          for (int r = 1; r < R; ++r) w[r] = cmul(w[r], twiddle(tw, r * k * step));
        }
      }
      DFT<R, DIR, T>::run(w);
    }
  }

  template <int MAPW, int MAPR, int Ns, int R, int Rn, class XLY = LY>
  __device__ __forceinline__ static void exchange(cx<T>* v, int t_in, T* lds) {
    constexpr int comps = SPLIT ? 2 : 1;
    constexpr bool local = wave_owner<MAPW, R>() != 0 && wave_owner<MAPW, R>() == wave_owner<MAPR, Rn>();
    static_assert(XLY::LP == LY::LP, "every exchange layout of a transform needs the same sequence stride");
    // opaque copy of the thread index: LDS addresses are recomputed here instead of
    // being hoisted (they are loop-invariant across transforms) and held live
    unsigned t = t_in;
    asm volatile("" : "+v"(t));
    __builtin_assume(t < (unsigned)NT);
#pragma unroll
    for (int c = 0; c < comps; ++c) {
      exchange_sync<local>();
#pragma unroll
      for (int b = 0; b < E / R; ++b) {
        unsigned s, j;
        bmap<MAPW, TS, NT, L, R, G>(t, b, s, j);
        const unsigned k = j % Ns;
        const unsigned base = (j - k) * R + k;
        const unsigned a0 = XLY::at(s, base);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const unsigned idx = XLY::linear(Ns) ? a0 + r * XLY::delta(Ns) : XLY::at(s, base + r * Ns);
          if (SPLIT) {
            lds[idx] = c == 0 ? v[b * R + r].x : v[b * R + r].y;
          } else {
            lds[2 * idx] = v[b * R + r].x;
            lds[2 * idx + 1] = v[b * R + r].y;
          }
        }
      }
      exchange_sync<local>();
#pragma unroll
      for (int b = 0; b < E / Rn; ++b) {
        unsigned s, j;
        bmap<MAPR, TS, NT, L, Rn, G>(t, b, s, j);
        const unsigned a0 = XLY::at(s, j);
#pragma unroll
        for (int r = 0; r < Rn; ++r) {
          const unsigned idx = XLY::linear(L / Rn) ? a0 + r * XLY::delta(L / Rn) : XLY::at(s, j + r * (L / Rn));
          if (SPLIT) {
            if (c == 0)
              v[b * Rn + r].x = lds[idx];
            else
              v[b * Rn + r].y = lds[idx];
          } else {
            v[b * Rn + r].x = lds[2 * idx];
            v[b * Rn + r].y = lds[2 * idx + 1];
          }
        }
      }
    }
  }

  template <int P, int NP, int Ns, int R>
  __device__ __forceinline__ static void passes(cx<T>* v, int t, T*, const cx<T>* tw, Plan<R>) {
    butterflies<map_of<P, NP>(), Ns, R>(v, t, tw);
  }
  template <int P, int NP, int Ns, int R, int Rn, int... Rs>
  __device__ __forceinline__ static void passes(cx<T>* v, int t, T* lds, const cx<T>* tw, Plan<R, Rn, Rs...>) {
    butterflies<map_of<P, NP>(), Ns, R>(v, t, tw);
    exchange<map_of<P, NP>(), map_of<P + 1, NP>(), Ns, R, Rn, XL<P>>(v, t, lds);
    // per-pass tables: pass P >= 1 takes Ns R entries, the next pass's start after them
    passes<P + 1, NP, Ns * R, Rn, Rs...>(v, t, lds, (PTAB && P >= 1) ? tw + Ns * R : tw, Plan<Rn, Rs...>{});
  }

  // all threads of the workgroup call run (it holds barriers)
  template <int... Rs>
  __device__ __forceinline__ static void run(cx<T>* v, int t, T* lds, const cx<T>* tw, Plan<Rs...> p) {
    static_assert(plan_prod<Rs...>() == L, "plan must multiply to L");
    passes<0, PlanAt<Rs...>::count, 1, Rs...>(v, t, lds, tw, p);
  }

  // run() after load_dft(): the first pass's butterflies are already applied
  template <int R0>
  __device__ __forceinline__ static void run_rest(cx<T>*, int, T*, const cx<T>*, Plan<R0>) {}
  template <int R0, int R1, int... Rs>
  __device__ __forceinline__ static void run_rest(cx<T>* v, int t, T* lds, const cx<T>* tw, Plan<R0, R1, Rs...>) {
    static_assert(plan_prod<R0, R1, Rs...>() == L, "plan must multiply to L");
    constexpr int NP = PlanAt<R0, R1, Rs...>::count;
    exchange<map_of<0, NP>(), map_of<1, NP>(), 1, R0, R1, XL<0>>(v, t, lds);
    passes<1, NP, R0, R1, Rs...>(v, t, lds, tw, Plan<R1, Rs...>{});
  }

  // Continue a transform whose first pass (radix R0, Ns = 1) the caller computed
  // and wrote to LDS itself (element (s, pos) at XL<0>::at(s, pos), followed by a
  // barrier): read pass 1's inputs with pass 1's mapping and run passes 1 ...
  template <int R0, int R1, int... Rs>
  __device__ __forceinline__ static void run_after_first(cx<T>* v, int t_in, T* lds, const cx<T>* tw,
                                                         Plan<R0, R1, Rs...>) {
    static_assert(plan_prod<R0, R1, Rs...>() == L, "plan must multiply to L");
    constexpr int NP = PlanAt<R0, R1, Rs...>::count;
    unsigned t = t_in;
    asm volatile("" : "+v"(t));
    __builtin_assume(t < (unsigned)NT);
#pragma unroll
    for (int b = 0; b < E / R1; ++b) {
      unsigned s, j;
      bmap<map_of<1, NP>(), TS, NT, L, R1, G>(t, b, s, j);
      const unsigned a0 = XL<0>::at(s, j);
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        const unsigned idx = XL<0>::linear(L / R1) ? a0 + r * XL<0>::delta(L / R1) : XL<0>::at(s, j + r * (L / R1));
        v[b * R1 + r].x = lds[2 * idx];
        v[b * R1 + r].y = lds[2 * idx + 1];
      }
    }
    passes<1, NP, R0, R1, Rs...>(v, t_in, lds, tw, Plan<R1, Rs...>{});
  }

  // first-pass load fused with the first pass's butterflies (which need no
  // twiddles), one butterfly group at a time, so only one group's loads are live
  template <int R0, class Ld>
  __device__ __forceinline__ static void load_dft(cx<T>* v, int t_in, Ld ld) {
    unsigned t = t_in;
    asm volatile("" : "+v"(t));
    __builtin_assume(t < (unsigned)NT);
#pragma unroll
    for (int b = 0; b < E / R0; ++b) {
      unsigned s, j;
      bmap<MAPF, TS, NT, L, R0, G>(t, b, s, j);
#pragma unroll
      for (int r = 0; r < R0; ++r) v[b * R0 + r] = ld(s, j + r * (L / R0));
      DFT<R0, DIR, T>::run(v + b * R0);
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  // first-pass load with the first pass's mapping: ld(s, pos)
  template <int R0, class Ld>
  __device__ __forceinline__ static void load(cx<T>* v, int t, Ld ld) {
#pragma unroll
    for (int b = 0; b < E / R0; ++b) {
      unsigned s, j;
      bmap<MAPF, TS, NT, L, R0, G>(t, b, s, j);
#pragma unroll
      for (int r = 0; r < R0; ++r) v[b * R0 + r] = ld(s, j + r * (L / R0));
    }
  }
  // last-pass store with the last pass's mapping: st(s, pos, value)
  template <int RL, class St>
  __device__ __forceinline__ static void store(const cx<T>* v, int t, St st) {
#pragma unroll
    for (int b = 0; b < E / RL; ++b) {
      unsigned s, j;
      bmap<MAPL, TS, NT, L, RL, G>(t, b, s, j);
#pragma unroll
      for (int r = 0; r < RL; ++r) st(s, j + r * (L / RL), v[b * RL + r]);
    }
  }
};


// Per-pass twiddle table of a plan (see BlockFFT2's PTAB): entries for passes
// P >= 1, Ns_P rows of R_P entries each.  Host side: the values are the W_L^e
// table's (the same rounding), conjugated for the inverse direction.
template <int R0>
constexpr int pass_table_rest(int, Plan<R0>) { return 0; }
template <int R0, int R1, int... Rs>
constexpr int pass_table_rest(int ns, Plan<R0, R1, Rs...>) {
  return ns * R0 * R1 + pass_table_rest(ns * R0, Plan<R1, Rs...>{});
}
template <int R0, int... Rs>
constexpr int pass_table_entries(Plan<R0, Rs...>) {
  return pass_table_rest(1, Plan<R0, Rs...>{});
}
// fill out[0 .. pass_table_entries) from the W_L^e table h (h[e] = exp(-2 pi i e / L))
template <typename T, int R0>
void fill_pass_rest(cx<T>*, const cx<T>*, int, int, int, Plan<R0>) {}
template <typename T, int R0, int R1, int... Rs>
void fill_pass_rest(cx<T>* out, const cx<T>* h, int L, int dir, int ns, Plan<R0, R1, Rs...>) {
  const int Ns = ns * R0, R = R1, step = L / (Ns * R);
  for (int k = 0; k < Ns; ++k)
    for (int r = 0; r < R; ++r) {
      cx<T> w{};
      if (r >= 1) {
        w = h[r * k * step];
        if (dir > 0) w.y = -w.y;
      }
      out[k * R + (r >= 1 ? r - 1 : R - 1)] = w;
    }
  fill_pass_rest(out + Ns * R, h, L, dir, Ns, Plan<R1, Rs...>{});
}
template <typename T, int R0, int... Rs>
void fill_pass_table(cx<T>* out, const cx<T>* h, int L, int dir, Plan<R0, Rs...>) {
  fill_pass_rest(out, h, L, dir, 1, Plan<R0, Rs...>{});
}
}  // namespace azb
