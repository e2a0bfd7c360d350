"""Pins the FP64 CPU oracle against brute force, closed forms and the reference's
own known-answer tests, so that a plausible mistake (dropped term, wrong sign or
index, transposed operand) fails at least one check.  CPU only."""
import math

import numpy as np
import pytest

from oracle import azeban_oracle as O
from azeban_amd import init as I


def signed_k(N):
    """2 pi times the signed wavenumber of index i: i - N for i >= N/2 + 1 (so N/2 stays
    positive, as in incompressible_euler.hpp:245-258 and leray.cpp:33-37)."""
    i = np.arange(N)
    return 2 * np.pi * np.where(i >= N // 2 + 1, i - N, i)


def dft_matrix(n, sign):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def brute_r2c(u):
    """Separable brute-force DFT, last axis first (test/azeban/operations/fft.cpp:35-59,146-150)."""
    n = u.shape[-1]
    out = u.astype(complex)
    for ax in range(u.ndim - 1, 0, -1):
        out = np.moveaxis(np.tensordot(dft_matrix(out.shape[ax], -1), np.moveaxis(out, ax, 0), axes=1), 0, ax)
    return out[..., : n // 2 + 1]


def brute_c2r(X, n):
    """Complex inverse DFT over the leading axes, then the real inverse of the last axis
    using Re(X_0), Re(X_{n/2}) and 2 Re(X_k e^{2 pi i k x/n}) for 0 < k < n/2."""
    Y = X.astype(complex)
    for ax in range(1, X.ndim - 1):
        Y = np.moveaxis(np.tensordot(dft_matrix(n, +1), np.moveaxis(Y, ax, 0), axes=1), 0, ax)
    x = np.arange(n)
    out = np.real(Y[..., 0])[..., None] + np.real(Y[..., n // 2])[..., None] * ((-1.0) ** x)
    for k in range(1, n // 2):
        out = out + 2 * np.real(Y[..., k][..., None] * np.exp(2j * np.pi * k * x / n))
    return out


# --------------------------------------------------------------------------- FFT
# workers: numpy's single-threaded pocketfft (None) or scipy.fft's threaded one
@pytest.mark.parametrize("workers", [None, 4])
@pytest.mark.parametrize("dim,n,nvars", [(1, 128, 1), (1, 128, 2), (2, 128, 1), (2, 128, 2),
                                         (3, 32, 1), (3, 32, 2), (1, 96, 2), (2, 48, 1)])
def test_r2c_matches_brute_force_dft(dim, n, nvars, workers):
    rng = np.random.default_rng(0)
    u = rng.uniform(-1, 1, (nvars,) + (n,) * dim)
    assert np.max(np.abs(O.fft_r2c(u, workers) - brute_r2c(u))) <= 1e-8


@pytest.mark.parametrize("workers", [None, 4])
@pytest.mark.parametrize("dim,n", [(1, 64), (2, 32), (3, 16), (2, 24)])
def test_c2r_non_hermitian_convention(dim, n, workers):
    rng = np.random.default_rng(1)
    shape = (2,) + (n,) * (dim - 1) + (n // 2 + 1,)
    X = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert np.max(np.abs(O.fft_c2r(X, n, workers) - brute_c2r(X, n))) <= 1e-9


@pytest.mark.parametrize("dim,N", [(2, 64), (3, 32)])
def test_dudt_threaded_ffts_agree(dim, N):
    """The RHS with scipy.fft's threaded transforms (used for 512^3 checks) equals the
    single-threaded numpy one to round-off."""
    rng = np.random.default_rng(4)
    uh = O.leray(O.fft_r2c(rng.standard_normal((dim,) + (N,) * dim)))
    visc = O.baseline_visc(N)
    a = O.IncompressibleEuler(O.Grid(dim, N), visc).dudt(uh)
    b = O.IncompressibleEuler(O.Grid(dim, N), visc, workers=4).dudt(uh)
    assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(a)


@pytest.mark.parametrize("dim,n", [(1, 128), (2, 128), (3, 32)])
def test_round_trip_scales_by_n_to_the_d(dim, n):
    rng = np.random.default_rng(2)
    u = rng.uniform(-1, 1, (2,) + (n,) * dim)
    assert np.max(np.abs(O.fft_c2r(O.fft_r2c(u), n) / n**dim - u)) <= 1e-12


def test_single_mode_closed_form():
    """cos(2 pi (i + j) / n) has exactly n^2/2 at (1, 1) and (n-1, ...) mirrored, 0 elsewhere."""
    n = 32
    i = np.arange(n)
    u = np.cos(2 * np.pi * (i[:, None] + i[None, :]) / n)[None]
    uh = O.fft_r2c(u)[0]
    expect = np.zeros_like(uh)
    expect[1, 1] = n * n / 2
    assert np.max(np.abs(uh - expect)) <= 1e-10
    # forward sign e^{-ikx}: sin(2 pi i/n) -> -i n/2 at k = 1
    s = np.sin(2 * np.pi * i / n)[None]
    assert abs(O.fft_r2c(s)[0, 1] - (-0.5j * n)) <= 1e-12


def test_next_good_size():
    def brute(nmin):
        n = nmin
        while True:
            m = n
            for p in (2, 3, 5, 7):
                while m % p == 0:
                    m //= p
            if m == 1:
                return n
            n += 1

    for nmin in list(range(1, 300)) + [97, 385, 769, 1537]:
        assert O.next_good_size(nmin) == brute(nmin)
    for N in (64, 128, 256, 512, 1024):
        assert O.next_good_size((3 * N + 1) // 2) == 3 * N // 2


# --------------------------------------------------------------------------- pad
def to_real_k(k, n):
    return k - n if k >= n // 2 + 1 else k


@pytest.mark.parametrize("dim,N,M", [(1, 16, 24), (1, 16, 32), (2, 16, 24), (2, 8, 16), (3, 8, 12), (2, 16, 16)])
def test_copy_to_padded_structural(dim, N, M):
    """test/azeban/operations/copy_to_padded.cpp:8-117: the padded array equals the
    source at the same signed wavenumber and is zero elsewhere."""
    rng = np.random.default_rng(3)
    src_shape = (N,) * (dim - 1) + (N // 2 + 1,)
    dst_shape = (M,) * (dim - 1) + (M // 2 + 1,)
    src = rng.uniform(-1, 1, src_shape) + 1j * rng.uniform(-1, 1, src_shape)
    dst = O.copy_to_padded(src, dst_shape)
    m = {}
    for idx in np.ndindex(*src_shape):
        key = tuple(to_real_k(i, N) for i in idx[:-1]) + (idx[-1],)
        m[key] = src[idx]
    for idx in np.ndindex(*dst_shape):
        key = tuple(to_real_k(i, M) for i in idx[:-1]) + (idx[-1],)
        assert dst[idx] == m.get(key, 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_copy_from_padded_round_trip_and_signed_modes(dim):
    """test/azeban/operations/copy_from_padded.cpp:21-66: unpadding a padded array gives
    the source back exactly, for M in {N, 3N/2, 2N}; and each kept entry is the padded
    entry at the same signed wavenumber (independent of copy_to_padded)."""
    rng = np.random.default_rng(5)
    for N in (8, 16) if dim == 3 else (8, 16, 32):
        for M in (N, 3 * N // 2, 2 * N):
            src_shape = (N,) * (dim - 1) + (N // 2 + 1,)
            dst_shape = (M,) * (dim - 1) + (M // 2 + 1,)
            src = rng.uniform(-1, 1, src_shape) + 1j * rng.uniform(-1, 1, src_shape)
            assert np.array_equal(O.copy_from_padded(O.copy_to_padded(src, dst_shape), src_shape), src)
            big = rng.uniform(-1, 1, dst_shape) + 1j * rng.uniform(-1, 1, dst_shape)
            small = O.copy_from_padded(big, src_shape)
            by_k = {tuple(to_real_k(i, M) for i in idx[:-1]) + (idx[-1],): big[idx]
                    for idx in np.ndindex(*dst_shape)}
            for idx in np.ndindex(*src_shape):
                assert small[idx] == by_k[tuple(to_real_k(i, N) for i in idx[:-1]) + (idx[-1],)]


# --------------------------------------------------------------------------- norm / SV
def test_max_norm_known_answer():
    n = 1000 * 1000
    assert abs(O.max_norm(np.full(n, float(n))) - n) <= 1e-10


@pytest.mark.parametrize("k0", ["1", "sqrt", "N"])
def test_smooth_cutoff_monotone_in_unit_interval(k0):
    """test/azeban/equations/spectral_viscosity.cpp:35-47 and the SmoothCutoff1D cases."""
    N = 16
    while N <= 4096:
        kc = {"1": 1.0, "sqrt": math.sqrt(N), "N": float(N)}[k0]
        v = O.SmoothCutoff1D(0.05, 1, kc)
        ks = 2 * math.pi * np.arange(N)
        q = v.Qk(ks)
        assert np.all((q >= 0) & (q <= 1))
        assert np.all(np.diff(q) >= 0)
        N *= 2


def test_smooth_cutoff_closed_form_and_factory():
    v = O.baseline_visc(128)
    assert v.eps == pytest.approx(0.05 / 128, rel=1e-15)
    assert v.k0 == pytest.approx(2 * math.pi * 128 / 4, rel=1e-15)
    k = np.linspace(0, 3 * v.k0, 101)
    expect = -v.eps * k**2 * (1 - np.exp(-((k / v.k0) ** 18)))
    assert np.allclose(v.eval(k), expect, rtol=1e-12, atol=0)
    # factory defaults: theta = (2s-1)/(2s), m0 = 2 pi (spectral_viscosity_factory.hpp:33-41)
    d = O.make_smooth_cutoff_1d(64, 0.05)
    assert d.k0 == pytest.approx(2 * math.pi * 64**0.5) and d.eps == pytest.approx(0.05 / 64)


# --------------------------------------------------------------------------- RHS
def eq_for(dim, N):
    return O.IncompressibleEuler(O.Grid(dim, N), O.baseline_visc(N))


def test_tg2d_is_steady():
    N = 64
    eq = eq_for(2, N)
    uh = O.initialize(I.taylor_green_2d(N))
    d = eq.dudt(uh)
    assert np.max(np.abs(d)) <= 1e-12 * np.max(np.abs(uh))
    assert eq.u_max == pytest.approx((N / (1.5 * N)) ** 2, rel=1e-14)
    assert eq.dt(0.2) == pytest.approx(0.2 / (N * (N / (1.5 * N)) ** 2), rel=1e-14)


def test_tg3d_nonlinear_term_closed_form():
    """For 3D Taylor-Green (x = 2 pi X), (u.grad)u = 2 pi (sin2x cos^2 z, sin2y cos^2 z, 0)/2,
    so dudt = -Leray(FFT of that); the viscous term is below round-off at this cutoff."""
    N = 32
    eq = eq_for(3, N)
    uh = O.initialize(I.taylor_green_3d(N))
    X, Y, Z = (2 * np.pi * a for a in np.meshgrid(*([np.arange(N) / N] * 3), indexing="ij"))
    adv = 2 * np.pi * np.stack([0.5 * np.sin(2 * X) * np.cos(Z) ** 2,
                                0.5 * np.sin(2 * Y) * np.cos(Z) ** 2, np.zeros_like(X)])
    expect = -O.leray(O.fft_r2c(adv))
    got = eq.dudt(uh)
    assert np.max(np.abs(got - expect)) <= 1e-9 * np.max(np.abs(expect))
    assert eq.u_max == pytest.approx((N / (1.5 * N)) ** 3, rel=1e-12)


def spectral_advection(uh, N):
    """(u.grad)u in physical space with spectral derivatives, then FFT and Leray."""
    dim = uh.shape[0]
    u = O.fft_c2r(uh, N) / N**dim
    k = [2 * np.pi * np.fft.fftfreq(N, 1.0 / N)] * (dim - 1) + [2 * np.pi * np.arange(N // 2 + 1)]
    K = np.meshgrid(*k, indexing="ij")
    adv = np.zeros_like(u)
    for i in range(dim):
        for j in range(dim):
            adv[i] += u[j] * O.fft_c2r((1j * K[j] * uh[i])[None], N)[0] / N**dim
    return -O.leray(O.fft_r2c(adv))


@pytest.mark.parametrize("dim,N", [(2, 32), (3, 16)])
def test_dudt_matches_advective_form_on_band_limited_field(dim, N):
    """A divergence-free field with |k_i| < N/4: products are exact on the N-grid, so the
    dealiased conservative RHS must equal the advective form computed independently."""
    rng = np.random.default_rng(4)
    shape = (dim,) + (N,) * (dim - 1) + (N // 2 + 1,)
    uh = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    kf = np.abs(np.fft.fftfreq(N, 1.0 / N))
    kl = np.arange(N // 2 + 1)
    mask = np.ones(shape[1:], bool)
    grids = np.meshgrid(*([kf] * (dim - 1) + [kl]), indexing="ij")
    for g in grids:
        mask &= g < N // 4
    uh = uh * mask
    uh = O.fft_r2c(O.fft_c2r(uh, N))  # make it an exact real field's spectrum
    uh = O.leray(uh)
    eq = O.IncompressibleEuler(O.Grid(dim, N), O.SmoothCutoff1D(0.0, 1.0, 1.0))
    got = eq.dudt(uh)
    expect = spectral_advection(uh, N)
    assert np.max(np.abs(got - expect)) <= 1e-10 * np.max(np.abs(expect))


def test_dudt_leray_and_zero_mode():
    N = 32
    eq = eq_for(2, N)
    uh = O.initialize(I.fractional_brownian(2, N, seed=2))
    d = eq.dudt(uh)
    assert np.all(d[:, 0, 0] == 0)
    kx = signed_k(N)[:, None]
    ky = 2 * np.pi * np.arange(N // 2 + 1)[None, :]
    assert np.max(np.abs(kx * d[0] + ky * d[1])) <= 1e-9 * np.max(np.abs(d))


# --------------------------------------------------------------------------- RK3
def test_rk3_linear_closed_form():
    """A shear flow u = (sin(2 pi m y), 0) has (u.grad)u = 0, so each step multiplies
    the mode by the SSP-RK3 stability polynomial R(nu dt) = 1 + z + z^2/2 + z^3/6."""
    N = 64
    m = 20  # above the cutoff N/4: nu is O(10)
    eq = eq_for(2, N)
    X, Y = np.meshgrid(np.arange(N) / N, np.arange(N) / N, indexing="ij")
    uh = O.initialize(np.stack([np.sin(2 * np.pi * m * Y), 0 * Y]))
    nu = eq.visc.eval(2 * np.pi * m)
    sim = O.Simulation(eq, uh, 0.2)
    sim.simulate_until(1.0, max_steps=6)
    ratio = np.prod([1 + nu * dt + (nu * dt) ** 2 / 2 + (nu * dt) ** 3 / 6 for dt in sim.dts])
    assert sim.u[0, 0, m] == pytest.approx(uh[0, 0, m] * ratio, rel=1e-12)
    # CFL step: u_max = max|u| / M^2 on the padded grid of the unnormalised transform,
    # where sin(2 pi m j / M) reaches 1 (m = 20, M = 96, j = 6), so u_max = (N/M)^2
    M = 3 * N // 2
    assert sim.dts[0] == pytest.approx(0.2 / (N * (N / M) ** 2), rel=1e-14)
    amp1 = 1 + nu * sim.dts[0] + (nu * sim.dts[0]) ** 2 / 2 + (nu * sim.dts[0]) ** 3 / 6
    assert sim.dts[1] == pytest.approx(0.2 / (N * (N / M) ** 2 * amp1), rel=1e-13)


def test_simulate_until_lands_on_t():
    N = 32
    eq = eq_for(2, N)
    sim = O.Simulation(eq, O.initialize(I.double_shear_layer(N)), 0.2)
    sim.simulate_until(0.05)
    assert sim.time == 0.05
    assert sum(sim.dts) == pytest.approx(0.05, rel=1e-14)
    # all but the last step take dt = C / (N u_max); the last is clamped onto t
    assert sim.dts[-1] <= sim.dts[0]


# --------------------------------------------------------------------------- leray / spectrum
@pytest.mark.parametrize("dim,N", [(2, 32), (3, 16)])
def test_leray_divergence_free_and_idempotent(dim, N):
    rng = np.random.default_rng(5)
    u = rng.standard_normal((dim,) + (N,) * dim)
    uh = O.fft_r2c(u)
    p = O.leray(uh)
    k = [signed_k(N)] * (dim - 1) + [2 * np.pi * np.arange(N // 2 + 1)]
    K = np.meshgrid(*k, indexing="ij")
    div = sum(K[d] * p[d] for d in range(dim))
    assert np.max(np.abs(div)) <= 1e-10 * np.max(np.abs(p))
    assert np.max(np.abs(O.leray(p) - p)) <= 1e-12 * np.max(np.abs(p))


@pytest.mark.parametrize("dim,N", [(2, 32), (3, 16)])
def test_energy_spectrum_parseval(dim, N):
    """sum_K E(K) = 1/2 <|u|^2> once every Nyquist index is zeroed (the reference's
    projection maps index N/2 to +N/2, which is not Hermitian-paired there)."""
    uh = O.initialize(I.fractional_brownian(dim, N, seed=1))
    for ax in range(1, dim + 1):
        idx = [slice(None)] * (dim + 1)
        idx[ax] = N // 2
        uh[tuple(idx)] = 0
    u = O.fft_c2r(uh, N) / N**dim
    e = O.energy_spectrum(uh, N)
    assert e.sum() == pytest.approx(0.5 * np.mean(np.sum(u * u, axis=0)), rel=1e-12)
    assert len(e) == N // 2 + 1


# Single-mode known answers of the reference's shell rule
# (include/azeban/operations/spectrum.hpp:33-118, energy_spectrum.cpp:10-50): a mode
# stored at row index i (signed k = i - N for i >= N/2 + 1, so index N/2 is +N/2) and
# last-axis index k (0 <= k <= N/2, never signed) lands in the L-infinity shell
# K = max(|k_i|), and counts 1/2 |u_hat / N^d|^2 once for k = 0 and twice (its
# conjugate partner) for every k != 0, the Nyquist index k = N/2 included.
# Each case: (index tuple, component, shell K, weight).  The shells are written out by
# hand; an L2 shell rint(|k|) differs on every oblique case ((-3, 3) -> 4, (2, -5, 1) -> 5
# vs 6, ...), and dropping the Nyquist doubling changes the weight of the k = N/2 cases.
SPECTRUM_CASES = {
    2: (16, [((13, 3), 0, 3, 2.0),     # k = (-3, 3): L-inf 3, L2 4
             ((5, 0), 1, 5, 1.0),      # k_last = 0: no conjugate partner counted
             ((2, 8), 0, 8, 2.0),      # k_last = N/2 (Nyquist) is doubled too
             ((8, 1), 1, 8, 2.0),      # row index N/2 is +N/2
             ((9, 2), 0, 7, 2.0),      # row index N/2 + 1 is -7
             ((15, 7), 1, 7, 2.0),     # k = (-1, 7): L2 7.07 -> 7 too, weight 2
             ((0, 0), 0, 0, 1.0)]),
    3: (16, [((2, 11, 1), 0, 5, 2.0),  # k = (2, -5, 1): L-inf 5, L2 rint(5.48) = 5
             ((12, 3, 0), 2, 4, 1.0),  # k = (-4, 3, 0): L2 5, no doubling
             ((1, 2, 8), 1, 8, 2.0),   # Nyquist k_last doubled
             ((8, 0, 0), 0, 8, 1.0),   # +N/2 along axis 0, k_last = 0
             ((13, 13, 3), 2, 3, 2.0),  # k = (-3, -3, 3): L2 rint(5.2) = 5
             ((0, 9, 6), 1, 7, 2.0),   # k = (0, -7, 6): L2 rint(9.2) = 9
             ((0, 0, 0), 1, 0, 1.0)]),
}


def _check_single_modes(spectrum, dim):
    """Every SPECTRUM_CASES entry through `spectrum(u_hat, N)`; raises AssertionError on
    the first mismatch."""
    N, cases = SPECTRUM_CASES[dim]
    a = 3.0 - 4.0j  # |a|^2 = 25
    for idx, comp, K, w in cases:
        uh = np.zeros((dim,) + (N,) * (dim - 1) + (N // 2 + 1,), dtype=np.complex128)
        uh[(comp,) + idx] = a * N**dim  # |u_hat / N^d|^2 = 25
        e = spectrum(uh, N)
        want = np.zeros(N // 2 + 1)
        want[K] = 0.5 * 25.0 * w
        assert np.array_equal(e, want), (idx, comp, K, w, e)


@pytest.mark.parametrize("dim", [2, 3])
def test_energy_spectrum_single_modes(dim):
    _check_single_modes(O.energy_spectrum, dim)


def test_energy_spectrum_single_modes_reject_other_shell_rules():
    """The single-mode cases fail an L2-shell, no-Nyquist-doubling spectrum (the
    variant the Parseval test cannot tell apart), so they pin the shell rule."""

    def l2_shells(uh, N):
        dim = uh.ndim - 1
        kf = O._signed(np.arange(N), N)
        kl = np.arange(uh.shape[-1])
        ks = np.meshgrid(*([kf] * (dim - 1) + [kl]), indexing="ij")
        K = np.rint(np.sqrt(sum(k.astype(float) ** 2 for k in ks))).astype(int)
        w = np.where((ks[-1] != 0) & (ks[-1] != N // 2), 2.0, 1.0)
        e = 0.5 * np.sum(np.abs(uh / float(N) ** dim) ** 2, axis=0) * w
        spec = np.zeros(N // 2 + 1)
        keep = K <= N // 2
        np.add.at(spec, K[keep], e[keep])
        return spec

    for dim in (2, 3):
        with pytest.raises(AssertionError):
            _check_single_modes(l2_shells, dim)


def test_energy_spectrum_of_physical_fields():
    """Interior modes carry the physical energy: u = cos(2 pi (3x - 2y)) has
    1/2 <u^2> = 1/4, all in shell 3; k_last = 0 splits over the two stored partners
    (u = cos(2 pi 5 x): 1/4 in shell 5).  The Nyquist line u = (-1)^j along the last
    axis has 1/2 <u^2> = 1/2 but is stored once and doubled, so the reference's rule
    reports 1 in shell N/2."""
    N = 16
    x = np.arange(N) / N
    X, Y = np.meshgrid(x, x, indexing="ij")
    for u, K, want in [(np.cos(2 * np.pi * (3 * X - 2 * Y)), 3, 0.25),
                       (np.cos(2 * np.pi * 5 * X), 5, 0.25),
                       (np.cos(np.pi * N * Y), N // 2, 1.0)]:
        field = np.stack([u, np.zeros_like(u)])
        e = O.energy_spectrum(O.fft_r2c(field), N)
        assert e[K] == pytest.approx(want, rel=1e-13)
        assert np.sum(np.abs(np.delete(e, K))) <= 1e-25


# --------------------------------------------------------------------------- golden
def test_golden_fixture_2d():
    """The committed fixture (tests/golden/make_golden.py) still matches the oracle."""
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "golden_2d_n64.npz")
    g = np.load(path)
    eq = eq_for(2, 64)
    assert np.max(np.abs(eq.dudt(g["u0"]) - g["dudt"])) <= 1e-10 * np.max(np.abs(g["dudt"]))
    sim = O.Simulation(eq, g["u0"], 0.2)
    sim.simulate_until(1.0, max_steps=5)
    assert np.max(np.abs(sim.u - g["u5"])) <= 1e-10 * np.max(np.abs(g["u5"]))
    assert np.allclose(sim.dts, g["dts"], rtol=1e-14, atol=0)


@pytest.mark.parametrize("dim,N", [(2, 64), (3, 32)])
def test_threaded_rows_give_identical_dudt(dim, N):
    """compute_B / compute_dudt split over row chunks on several threads (the reference's
    OpenMP loops, used by the CPU baseline) give exactly the single-threaded result."""
    u = O.initialize(I.fractional_brownian(dim, N, seed=3))
    a = O.IncompressibleEuler(O.Grid(dim, N), O.baseline_visc(N)).dudt(u)
    for threads in (2, 5):
        b = O.IncompressibleEuler(O.Grid(dim, N), O.baseline_visc(N), threads=threads).dudt(u)
        assert np.array_equal(a, b), threads
