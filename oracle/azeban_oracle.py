"""FP64 CPU oracle for azeban's spectral-viscosity Euler RHS and SSP-RK3 stepper.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path (``azeban_amd``) never
calls into it.

This is a plain numpy restatement of the reference algorithm (TobiasRohner/azeban),
function by function, each citing the reference file:line it follows.  The DFT
engine is ``numpy.fft`` (pocketfft), standing in for FFTW 3.3.x (pinned by the
reference's ``conanfile.txt`` as fftw/3.3.9): FFTW computes the unnormalised
DFT with e^{-ikx} forward / e^{+ikx} backward, and its multi-dimensional c2r does
the complex passes over the leading axes first and the 1-D c2r over the last axis
last, discarding Im of the k=0 and k=M/2 entries of that axis.  numpy's
``irfftn`` has exactly that structure; tests/test_oracle.py checks both the DFT
and that convention against brute force.

Parity status per function (what pins it; see tests/test_oracle.py):
  next_good_size          pinned: brute-force search over 2^a3^b5^c7^d
  copy_to_padded          pinned: reference structural test rules (exact)
  fft_r2c / fft_c2r       pinned: brute-force DFT (reference fft test), round trip,
                          single-mode closed forms, non-Hermitian k_last=0 rule
  max_norm                pinned: reference known-answer (constant array)
  SmoothCutoff1D          pinned: reference monotonicity/[0,1] property + closed form
  compute_B / dudt        pinned: 2D/3D Taylor-Green closed forms (steady state /
                          analytic nonlinear term), Leray invariant k.dudt = 0,
                          dudt(k=0)=0, a brute-force physical-space check of
                          u.grad(u) on a dealiased field
  ssp_rk3 / simulate      pinned: linear-ODE closed form (RK3 stability polynomial),
                          dt sequence and last-step clamp
  leray                   pinned: divergence at round-off, idempotence
  energy_spectrum         pinned: single-mode known answers (L-infinity shell index
                          and conjugate doubling at k_last = 0, interior, N/2, which
                          an L2-shell or no-Nyquist-doubling rule fails), physical-
                          field energies, Parseval identity
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

PI = math.pi


# ---------------------------------------------------------------------------
# Grid and padded size  (include/azeban/grid.hpp:41-53, src/azeban/operations/fft.cpp:27-60)
# ---------------------------------------------------------------------------
def next_good_size(nmin: int) -> int:
    """Smallest 2^a 3^b 5^c 7^d >= nmin (src/azeban/operations/fft.cpp:27-60; the reference
    collects one candidate per (p3, p5, p7) and takes the minimum, which for every
    n = 3*2^k it is asked about equals this first-fit search)."""
    n = max(1, int(nmin))
    while True:
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


@dataclass
class Grid:
    """Grid<Dim> (include/azeban/grid.hpp:41-53): N_phys, N_phys_pad = 3N/2."""

    dim: int
    N: int
    M: int = 0

    def __post_init__(self):
        if self.M == 0:
            self.M = 3 * self.N // 2  # grid.hpp:45 / grid_factory.hpp:33
        self.Nf = self.N // 2 + 1
        self.Mf = self.M // 2 + 1

    def shape_fourier(self, n_vars):  # grid.hpp:126-133
        return (n_vars,) + (self.N,) * (self.dim - 1) + (self.Nf,)

    def shape_fourier_pad(self, n_vars):  # grid.hpp:157-165
        return (n_vars,) + (self.M,) * (self.dim - 1) + (self.Mf,)

    def shape_phys(self, n_vars):
        return (n_vars,) + (self.N,) * self.dim

    def shape_phys_pad(self, n_vars):
        return (n_vars,) + (self.M,) * self.dim


# ---------------------------------------------------------------------------
# Spectral viscosity (include/azeban/equations/spectral_viscosity.hpp:27-38,56-77)
# ---------------------------------------------------------------------------
@dataclass
class SmoothCutoff1D:
    eps: float
    s: float
    k0: float

    def Qk(self, k):
        """1 - exp(-(|k|/k0)^18) by repeated squaring (spectral_viscosity.hpp:64-72)."""
        k1 = np.abs(k) / self.k0
        k2 = k1 * k1
        k4 = k2 * k2
        k8 = k4 * k4
        k16 = k8 * k8
        k18 = k16 * k2
        return 1 - np.exp(-k18)

    def eval(self, k):
        """nu(k) = -eps k^{2s} Q(k)  (spectral_viscosity.hpp:31-34)."""
        return -self.eps * np.power(k, 2 * self.s) * self.Qk(k)


def make_smooth_cutoff_1d(N: int, eps: float, s: float = 1.0, theta=None, m0=None):
    """SmoothCutoff factory (include/azeban/equations/spectral_viscosity_factory.hpp:29-54)."""
    if theta is None:
        theta = (2 * s - 1) / (2 * s)
    if m0 is None:
        m0 = 2 * PI
    epsN = eps / float(N) ** (2 * s - 1)
    mN = m0 * float(N) ** theta
    return SmoothCutoff1D(epsN, s, mN)


def baseline_visc(N: int) -> SmoothCutoff1D:
    """eps = 0.05/N with the cutoff at wavenumber index N/4: SmoothCutoff1D(0.05/N, 1, 2*pi*N/4)."""
    return make_smooth_cutoff_1d(N, 0.05, 1.0, theta=1.0, m0=PI / 2)


# ---------------------------------------------------------------------------
# Zero padding (include/azeban/operations/copy_to_padded_functions.hpp:11-63,
#               src/azeban/operations/copy_to_padded.cpp:74-121)
# ---------------------------------------------------------------------------
def _full_dim_src(i, src, dst):
    """pad_full_dim: source index of padded index i, or -1 for a pad slot (:27-44)."""
    i = np.asarray(i)
    out = np.where(i < src // 2 + 1, i, i + src - dst)
    return np.where((i >= src // 2 + 1) & (i < src // 2 + 1 + dst - src), -1, out)


def copy_to_padded(src: np.ndarray, dst_shape, pad_value=0.0) -> np.ndarray:
    """One component: full (non-compact) axes keep i<=N/2 in place, move i>=N/2+1 to i+M-N,
    zero the gap; the compact last axis keeps k < src_shape and zeros the rest."""
    dst = np.full(dst_shape, pad_value, dtype=np.complex128)
    d = src.ndim
    idx = []
    for ax in range(d):
        n_dst = dst_shape[ax]
        n_src = src.shape[ax]
        ii = np.arange(n_dst)
        if ax == d - 1:  # compact dim (pad_compact_dim, :11-24)
            s = np.where(ii < n_src, ii, -1)
        else:
            s = _full_dim_src(ii, n_src, n_dst)
        idx.append(s)
    grids = np.meshgrid(*idx, indexing="ij")
    valid = np.ones(dst_shape, dtype=bool)
    for g in grids:
        valid &= g >= 0
    src_idx = tuple(np.where(valid, g, 0) for g in grids)
    dst[valid] = src[src_idx][valid]
    return dst



def copy_from_padded(src: np.ndarray, dst_shape) -> np.ndarray:
    """One component, the truncation inverse of copy_to_padded
    (include/azeban/operations/copy_from_padded_functions.hpp:11-27, unpad_full_dim /
    unpad_compact_dim; src/azeban/operations/copy_from_padded.cpp:28-86): full axes read
    i < N/2+1 from i and i >= N/2+1 from i + M - N; the compact last axis reads k from k."""
    d = src.ndim
    idx = []
    for ax in range(d):
        n_dst = dst_shape[ax]
        ii = np.arange(n_dst)
        if ax == d - 1:
            idx.append(ii)
        else:
            idx.append(np.where(ii < n_dst // 2 + 1, ii, ii + src.shape[ax] - n_dst))
    return src[np.ix_(*idx)].astype(np.complex128)

# ---------------------------------------------------------------------------
# FFT (src/azeban/operations/fftwfft.cpp:258-350): unnormalised, batched over
# the leading component axis, rank = dim.
# ---------------------------------------------------------------------------
def fft_r2c(u: np.ndarray, workers: int | None = None) -> np.ndarray:
    """FFTWFFT_R2C::forward: u (c, n, ..., n) real -> (c, n, ..., n/2+1), e^{-ikx}.
    workers: run scipy.fft's multithreaded pocketfft (the same algorithm as numpy's)
    instead of numpy's single-threaded one, for grids numpy takes minutes on."""
    axes = tuple(range(1, u.ndim))
    if workers:
        import scipy.fft

        return scipy.fft.rfftn(u, axes=axes, workers=workers)
    return np.fft.rfftn(u, axes=axes)


def fft_c2r(u_hat: np.ndarray, n: int, workers: int | None = None) -> np.ndarray:
    """FFTWFFT_R2C::backward: unnormalised, e^{+ikx}; complex passes over leading axes,
    then c2r over the last axis (which drops Im at k_last in {0, n/2})."""
    axes = tuple(range(1, u_hat.ndim))
    s = (n,) * (u_hat.ndim - 1)
    if workers:
        import scipy.fft

        return scipy.fft.irfftn(u_hat, s=s, axes=axes, workers=workers) * float(n) ** (u_hat.ndim - 1)
    return np.fft.irfftn(u_hat, s=s, axes=axes) * float(n) ** (u_hat.ndim - 1)


# ---------------------------------------------------------------------------
# Norm and axpby (include/azeban/operations/norm.hpp:72-91, axpby.hpp:28-56)
# ---------------------------------------------------------------------------
def max_norm(x: np.ndarray) -> float:
    return float(np.max(np.abs(x)))


def axpby(a, x, b, y):
    """y <- a x + b y (axpby.hpp:41-46)."""
    return a * x + b * y


# ---------------------------------------------------------------------------
# Incompressible Euler RHS (include/azeban/equations/incompressible_euler.hpp)
# ---------------------------------------------------------------------------
def _signed(i, N):
    return np.where(i >= N // 2 + 1, i - N, i)


def _pad_index(i, N, M):
    return np.where(i >= N // 2 + 1, M - N + i, i)


class IncompressibleEuler:
    """IncompressibleEuler<Dim, SmoothCutoff1D, NoForcing> (incompressible_euler.hpp:39-156)."""

    def __init__(self, grid: Grid, visc: SmoothCutoff1D, workers: int | None = None, threads: int = 1):
        self.grid = grid
        self.visc = visc
        self.workers = workers  # FFT threads (None: numpy's single-threaded pocketfft)
        # threads of compute_B / compute_dudt, which the reference runs as OpenMP loops
        # (incompressible_euler.hpp:176,193,236,290) while its FFTW stays single-threaded;
        # chunks of rows along the first axis, identical results for any count
        self.threads = max(1, int(threads))
        self.u_max = 0.0
        self.dim = grid.dim

    def _rows(self, n, fn):
        """fn(rows) over row chunks of the first spatial axis, on self.threads threads
        (numpy releases the interpreter lock inside its element-wise loops)."""
        if self.threads == 1:
            return [fn(slice(0, n))]
        from concurrent.futures import ThreadPoolExecutor

        step = -(-n // self.threads)
        chunks = [slice(a, min(n, a + step)) for a in range(0, n, step)]
        with ThreadPoolExecutor(self.threads) as ex:
            return list(ex.map(fn, chunks))

    def n_vars(self):
        return self.dim

    # incompressible_euler.hpp:103-120
    def dudt(self, u_hat: np.ndarray) -> np.ndarray:
        g = self.grid
        pad_shape = g.shape_fourier_pad(1)[1:]
        u_hat_pad = np.stack([copy_to_padded(u_hat[c], pad_shape) for c in range(self.dim)])
        u = fft_c2r(u_hat_pad, g.M, self.workers)
        del u_hat_pad  # intermediates are dropped as soon as used (large grids)
        self.u_max = max_norm(u) / float(g.M) ** self.dim  # :116
        B = self.compute_B(u)
        del u
        B_hat = fft_r2c(B, self.workers)
        del B
        return self.compute_dudt(B_hat, u_hat)

    def dt(self, C: float) -> float:
        """C / (N_phys * u_max)  (incompressible_euler.hpp:122-124)."""
        return C / (self.grid.N * self.u_max)

    def compute_B(self, u: np.ndarray) -> np.ndarray:
        """incompressible_euler.hpp:168-212 with incompressible_euler_functions.hpp:26-45."""
        g = self.grid
        norm = 1.0 / (math.sqrt(g.N) ** self.dim * math.sqrt(g.M) ** self.dim)
        nB = 3 if self.dim == 2 else 6
        B = np.empty((nB,) + u.shape[1:], dtype=u.dtype)

        def rows(r):
            if self.dim == 2:
                u1 = norm * u[0, r]
                u2 = norm * u[1, r]
                B[0, r], B[1, r], B[2, r] = u1 * u1, u1 * u2, u2 * u2
                return
            u1 = norm * u[0, r]
            u2 = norm * u[1, r]
            u3 = norm * u[2, r]
            B[0, r], B[1, r], B[2, r] = u1 * u1, u2 * u1, u2 * u2
            B[3, r], B[4, r], B[5, r] = u3 * u1, u3 * u2, u3 * u3

        self._rows(u.shape[1], rows)
        return B

    def wavenumbers(self):
        """Signed wavenumbers k = 2 pi i_ of the N-grid modes and the B_hat row indices."""
        g = self.grid
        i = np.arange(g.N)
        k_full = 2 * PI * _signed(i, g.N)
        iB = _pad_index(i, g.N, g.M)
        k_last = 2 * PI * np.arange(g.Nf)
        return k_full, iB, k_last

    def compute_dudt(self, B_hat: np.ndarray, u_hat: np.ndarray) -> np.ndarray:
        """computeDudt_cpu_{2d,3d} (incompressible_euler.hpp:231-347) with
        incompressible_euler_{2d,3d}_compute_L (incompressible_euler_functions.hpp:47-112)."""
        out = np.empty_like(u_hat)

        def rows(r):
            out[:, r] = self._compute_dudt_rows(B_hat, u_hat, r)

        self._rows(u_hat.shape[1], rows)
        return out

    def _compute_dudt_rows(self, B_hat: np.ndarray, u_hat_all: np.ndarray, r: slice) -> np.ndarray:
        """compute_dudt for the modes whose first index lies in r."""
        g = self.grid
        k_full, iB, k_last = self.wavenumbers()
        k_rows, iB_rows = k_full[r], iB[r]
        u_hat = u_hat_all[:, r]
        with np.errstate(divide="ignore", invalid="ignore"):
            if self.dim == 2:
                Bh = B_hat[:, iB_rows, : g.Nf]
                k1 = k_rows[:, None]
                k2 = k_last[None, :]
                absk2 = k1 * k1 + k2 * k2
                B11, B12, B22 = Bh[0], Bh[1], Bh[2]
                b1 = (1j * k1) * B11 + (1j * k2) * B12
                b2 = (1j * k1) * B12 + (1j * k2) * B22
                inv = 1.0 / absk2
                k11 = k1 * k1 * inv
                k12 = k1 * k2 * inv
                k22 = k2 * k2 * inv
                L1 = (1.0 - k11) * b1 + (0.0 - k12) * b2
                L2 = (0.0 - k12) * b1 + (1.0 - k22) * b2
                nu = self.visc.eval(np.sqrt(absk2))
                out = np.stack([-L1 + nu * u_hat[0], -L2 + nu * u_hat[1]])
            else:
                Bh = B_hat[:, iB_rows][:, :, iB, : g.Nf]
                k1 = k_rows[:, None, None]
                k2 = k_full[None, :, None]
                k3 = k_last[None, None, :]
                absk2 = k1 * k1 + k2 * k2 + k3 * k3
                B11, B21, B22, B31, B32, B33 = (Bh[m] for m in range(6))
                b1 = (1j * k1) * B11 + (1j * k2) * B21 + (1j * k3) * B31
                b2 = (1j * k1) * B21 + (1j * k2) * B22 + (1j * k3) * B32
                b3 = (1j * k1) * B31 + (1j * k2) * B32 + (1j * k3) * B33
                inv = 1.0 / absk2
                k11 = k1 * k1 * inv
                k12 = k1 * k2 * inv
                k13 = k1 * k3 * inv
                k22 = k2 * k2 * inv
                k23 = k2 * k3 * inv
                k33 = k3 * k3 * inv
                L1 = (1.0 - k11) * b1 + (0.0 - k12) * b2 + (0.0 - k13) * b3
                L2 = (0.0 - k12) * b1 + (1.0 - k22) * b2 + (0.0 - k23) * b3
                L3 = (0.0 - k13) * b1 + (0.0 - k23) * b2 + (1.0 - k33) * b3
                nu = self.visc.eval(np.sqrt(absk2))
                out = np.stack(
                    [-L1 + nu * u_hat[0], -L2 + nu * u_hat[1], -L3 + nu * u_hat[2]]
                )
        out[:, absk2 == 0] = 0
        return out


# ---------------------------------------------------------------------------
# SSP-RK3 and the simulation loop (include/azeban/evolution/ssp_rk3.hpp:48-64,
# src/azeban/simulation.cpp:89-102)
# ---------------------------------------------------------------------------
def ssp_rk3_integrate(eq: IncompressibleEuler, u: np.ndarray, max_dt: float, C: float):
    """Returns (u_new, dt); same axpby sequence as ssp_rk3.hpp:54-62."""
    u1 = eq.dudt(u)
    dt = min(eq.dt(C), max_dt)
    u1 = axpby(1.0, u, dt, u1)
    u2 = eq.dudt(u1)
    u2 = axpby(1.0, u1, dt, u2)
    u2 = axpby(3.0 / 4, u, 1.0 / 4, u2)
    u1 = eq.dudt(u2)
    u1 = axpby(1.0, u2, dt, u1)
    u = axpby(2.0 / 3, u1, 1.0 / 3, u)
    return u, dt


class Simulation:
    """Simulation<Dim> (include/azeban/simulation.hpp:30-93, src/azeban/simulation.cpp:89-102)."""

    def __init__(self, eq: IncompressibleEuler, u_hat: np.ndarray, C: float):
        self.eq = eq
        self.u = np.array(u_hat, dtype=np.complex128, copy=True)
        self.C = C
        self.time = 0.0
        self.dts = []

    def max_dt_visc(self):
        eps = self.eq.visc.eps
        return self.C * 2.0 / (eps * (PI * self.eq.grid.N) ** 2)

    def step(self, t: float):
        max_dt = min(t - self.time, self.max_dt_visc())
        self.u, dt = ssp_rk3_integrate(self.eq, self.u, max_dt, self.C)
        self.dts.append(dt)
        self.time += dt
        return dt

    def simulate_until(self, t: float, max_steps=None):
        n = 0
        while self.time < t:
            self.step(t)
            n += 1
            if max_steps is not None and n >= max_steps:
                return
        self.time = t


# ---------------------------------------------------------------------------
# Leray projection, initial data, energy spectrum
# ---------------------------------------------------------------------------
def leray(u_hat: np.ndarray) -> np.ndarray:
    """src/azeban/operations/leray.cpp:28-116 (k=0 mode set to 0)."""
    dim = u_hat.ndim - 1
    N = u_hat.shape[1]
    k_full = 2 * PI * _signed(np.arange(N), N)
    k_last = 2 * PI * np.arange(u_hat.shape[-1])
    out = np.empty_like(u_hat)
    with np.errstate(divide="ignore", invalid="ignore"):
        if dim == 2:
            k1 = k_full[:, None]
            k2 = k_last[None, :]
            absk2 = k1 * k1 + k2 * k2
            u1, u2 = u_hat[0], u_hat[1]
            out[0] = (1.0 - (k1 * k1) / absk2) * u1 + (0.0 - (k1 * k2) / absk2) * u2
            out[1] = (0.0 - (k2 * k1) / absk2) * u1 + (1.0 - (k2 * k2) / absk2) * u2
        else:
            k1 = k_full[:, None, None]
            k2 = k_full[None, :, None]
            k3 = k_last[None, None, :]
            absk2 = k1 * k1 + k2 * k2 + k3 * k3
            u1, u2, u3 = u_hat[0], u_hat[1], u_hat[2]
            out[0] = ((1.0 - (k1 * k1) / absk2) * u1 + (0.0 - (k1 * k2) / absk2) * u2
                      + (0.0 - (k1 * k3) / absk2) * u3)
            out[1] = ((0.0 - (k2 * k1) / absk2) * u1 + (1.0 - (k2 * k2) / absk2) * u2
                      + (0.0 - (k2 * k3) / absk2) * u3)
            out[2] = ((0.0 - (k3 * k1) / absk2) * u1 + (0.0 - (k3 * k2) / absk2) * u2
                      + (1.0 - (k3 * k3) / absk2) * u3)
    out[:, absk2 == 0] = 0
    return out


def initialize(u_phys: np.ndarray) -> np.ndarray:
    """Initializer::initialize: u_hat = FFT(u) unnormalised, then Leray
    (include/azeban/init/initializer.hpp:68-75)."""
    return leray(fft_r2c(u_phys))


def energy_spectrum(u_hat: np.ndarray, N: int) -> np.ndarray:
    """E(K) over L-infinity shells K = max|k_i|, conjugate doubling for k_last != 0
    (src/azeban/operations/energy_spectrum.cpp:10-50, include/azeban/operations/spectrum.hpp:33-118)."""
    dim = u_hat.ndim - 1
    norm = 1.0 / float(N) ** dim
    kf = np.abs(_signed(np.arange(N), N))
    kl = np.arange(u_hat.shape[-1])
    if dim == 2:
        K = np.maximum(kf[:, None], kl[None, :])
        w = np.where(kl[None, :] != 0, 2.0, 1.0) * np.ones_like(K, dtype=float)
    else:
        K = np.maximum(np.maximum(kf[:, None, None], kf[None, :, None]), kl[None, None, :])
        w = np.where(kl[None, None, :] != 0, 2.0, 1.0) * np.ones_like(K, dtype=float)
    e = 0.5 * np.sum(np.abs(norm * u_hat) ** 2, axis=0) * w
    spec = np.zeros(N // 2 + 1)
    np.add.at(spec, K.ravel(), e.ravel())
    return spec


def l2_rel(a: np.ndarray, b: np.ndarray) -> float:
    """Relative L2 of two half-spectra (conjugate-doubled like test/azeban/utils.hpp:61-111)."""
    w = np.ones(a.shape[-1])
    w[1:] = 2.0
    num = np.sum(np.abs(a - b) ** 2 * w)
    den = np.sum(np.abs(b) ** 2 * w)
    return math.sqrt(num / den) if den > 0 else math.sqrt(num)
