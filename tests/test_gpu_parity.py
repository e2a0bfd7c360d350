"""Parity of the HIP path (through the C ABI) against the FP64 CPU oracle.

Bars: FFT, RHS and single steps to round-off (relative 1e-12 or tighter);
trajectories to relative L2 <= 1e-11 after 100 steps (FP64); FP32-mixed to
relative L2 <= 1e-5; zero padding, Leray and max-norm known answers exactly.
"""
import math

import numpy as np
import pytest
import torch

from oracle import azeban_oracle as O

import azeban_amd as A
from azeban_amd import init as I

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def spectral_divergence(uh):
    """max |k . u_hat| / max |k| |u_hat| over the stored modes (signed wavenumbers on
    the leading axes, i - N for i >= N/2 + 1, as leray.cpp:33-37)."""
    dim = uh.shape[0]
    N = uh.shape[1]
    i = np.arange(N)
    ks = np.where(i >= N // 2 + 1, i - N, i).astype(float)
    kl = np.arange(uh.shape[-1], dtype=float)
    grids = np.meshgrid(*([ks] * (dim - 1) + [kl]), indexing="ij")
    div = sum(grids[c] * uh[c] for c in range(dim))
    kabs = np.sqrt(sum(g * g for g in grids))
    scale = float(np.max(kabs * np.sqrt(sum(np.abs(uh[c]) ** 2 for c in range(dim)))))
    return float(np.max(np.abs(div))) / scale


def random_state(dim, N, seed=3):
    print(f"[case] random fBm state dim={dim} N={N} seed={seed}", flush=True)
    return O.initialize(I.fractional_brownian(dim, N, seed=seed))


# ---------------------------------------------------------------------------
# FFT seams (test/azeban/operations/fft.cpp:219-354 analogue: n = 128, 1-2 vars)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dim,n,howmany", [(1, 128, 1), (1, 128, 2), (2, 128, 1), (2, 128, 2),
                                           (3, 128, 1), (3, 128, 2), (1, 96, 3), (1, 768, 2),
                                           (2, 192, 2), (3, 96, 1), (2, 1536, 1)])
def test_fft_r2c_c2r_fp64(cuda, dim, n, howmany):
    print(f"[case] fft dim={dim} n={n} howmany={howmany}", flush=True)
    rng = np.random.default_rng(7)
    u = rng.uniform(-1, 1, (howmany,) + (n,) * dim)
    ref = O.fft_r2c(u)
    got = host(A.fft_r2c(dev(u)))
    assert np.max(np.abs(got - ref)) <= 1e-8
    assert rel(got, ref) <= 1e-14
    back = host(A.fft_c2r(dev(ref), n))
    assert np.max(np.abs(back / n**dim - u)) <= 1e-8
    assert rel(back, O.fft_c2r(ref, n)) <= 1e-14


@pytest.mark.parametrize("dim,n", [(1, 128), (2, 128), (3, 96)])
def test_fft_c2r_non_hermitian_convention(cuda, dim, n):
    """A half spectrum whose k_last = 0 (and n/2) entries are not Hermitian: the
    result must equal complex passes on the leading axes, then a c2r that ignores
    Im at k_last in {0, n/2} (the FFTW/cuFFT convention the oracle follows)."""
    print(f"[case] c2r non-Hermitian dim={dim} n={n}", flush=True)
    rng = np.random.default_rng(11)
    shape = (2,) + (n,) * (dim - 1) + (n // 2 + 1,)
    X = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = host(A.fft_c2r(dev(X), n))
    assert rel(got, O.fft_c2r(X, n)) <= 1e-14


def test_fft_fp32(cuda):
    print("[case] fft fp32 3D n=96", flush=True)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, (2, 96, 96, 96))
    got = host(A.fft_r2c(dev(u.astype(np.float32))))
    assert rel(got, O.fft_r2c(u)) <= 1e-6


# ---------------------------------------------------------------------------
# pad / axpby / max_norm / leray / spectrum
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dim,N", [(2, 64), (2, 128), (3, 64)])
def test_copy_to_padded_exact(cuda, dim, N):
    print(f"[case] pad dim={dim} N={N}", flush=True)
    M = 3 * N // 2
    rng = np.random.default_rng(1)
    shape = (N,) * (dim - 1) + (N // 2 + 1,)
    src = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    got = host(A.copy_to_padded(dev(src), M))
    ref = O.copy_to_padded(src, (M,) * (dim - 1) + (M // 2 + 1,))
    assert np.array_equal(got, ref)


def test_max_norm_known_answer(cuda):
    """test/azeban/operations/norm.cpp:25-41: max_norm of a constant array of length n is n."""
    n = 1000 * 1000
    x = torch.full((n,), float(n), dtype=torch.float64, device="cuda")
    assert abs(A.max_norm(x) - n) <= 1e-10
    y = torch.linspace(-3.0, 2.0, 12345, dtype=torch.float64, device="cuda")
    assert A.max_norm(y) == 3.0


def test_axpby(cuda):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(1000) + 1j * rng.standard_normal(1000)
    y = rng.standard_normal(1000) + 1j * rng.standard_normal(1000)
    got = host(A.axpby_(0.75, dev(x), 0.25, dev(y)))
    assert np.allclose(got, 0.75 * x + 0.25 * y, rtol=0, atol=1e-15)


@pytest.mark.parametrize("dim,N", [(2, 128), (3, 64)])
def test_initialize_and_leray(cuda, dim, N):
    print(f"[case] initialize dim={dim} N={N}", flush=True)
    u = I.fractional_brownian(dim, N, seed=4)
    got = host(A.initialize(u))
    ref = O.initialize(u)
    assert rel(got, ref) <= 1e-13
    # divergence at round-off
    i = np.arange(N)
    ks = 2 * math.pi * np.where(i >= N // 2 + 1, i - N, i)
    k = [ks] * (dim - 1) + [2 * math.pi * np.arange(N // 2 + 1)]
    g = np.meshgrid(*k, indexing="ij")
    div = sum(g[d] * got[d] for d in range(dim))
    assert np.max(np.abs(div)) <= 1e-12 * np.max(np.abs(got)) * N


@pytest.mark.parametrize("dim,N", [(2, 128), (3, 64)])
def test_energy_spectrum(cuda, dim, N):
    uh = random_state(dim, N, seed=9)
    got = np.array(A.energy_spectrum(dev(uh)))
    ref = O.energy_spectrum(uh, N)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("dim", [2, 3])
def test_energy_spectrum_single_modes_gpu(cuda, dim):
    """The device shell sums on the hand-written single-mode cases of the reference's
    rule (L-infinity shells, k_last != 0 doubled including the Nyquist index): the
    same known answers the oracle is pinned with, exactly."""
    from test_oracle import _check_single_modes

    print(f"[case] energy spectrum single modes dim={dim}", flush=True)
    _check_single_modes(lambda uh, N: np.array(A.energy_spectrum(dev(uh))), dim)


def test_golden_fixture_2d_gpu(cuda):
    """The committed fixture (written by tests/golden/make_golden.py from the oracle):
    the HIP RHS of its state and its 5-step SSP-RK3 trajectory and dt sequence."""
    import os

    print("[case] golden fixture 2D N=64", flush=True)
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_2d_n64.npz"))
    N = 64
    g, eq, _ = make_pair(2, N)
    u = dev(fx["u0"])
    d = torch.empty_like(u)
    eq.dudt(d, u)
    assert rel(host(d), fx["dudt"]) <= 1e-12
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(2), eq), u)
    dts = []
    for _ in range(5):
        t0 = sim.time()
        sim.simulate_until(1.0, max_steps=1, chunk=1)
        dts.append(sim.time() - t0)
    assert rel(host(sim.u()), fx["u5"]) <= 1e-12
    assert np.allclose(dts, fx["dts"], rtol=1e-12, atol=0)


# ---------------------------------------------------------------------------
# RHS (IncompressibleEuler::dudt)
# ---------------------------------------------------------------------------
def make_pair(dim, N, precision="fp64"):
    g = A.Grid(dim, N)
    visc = A.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4)
    eq = A.IncompressibleEuler(g, visc, precision)
    oeq = O.IncompressibleEuler(O.Grid(dim, N), O.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4))
    return g, eq, oeq


@pytest.mark.parametrize("dim,N", [(2, 64), (2, 128), (2, 256), (2, 512), (2, 1024), (3, 64), (3, 128)])
def test_dudt_fp64(cuda, dim, N):
    uh = random_state(dim, N)
    g, eq, oeq = make_pair(dim, N)
    ref = oeq.dudt(uh)
    u = dev(uh)
    out = torch.empty_like(u)
    eq.dudt(out, u)
    got = host(out)
    assert rel(got, ref) <= 1e-12, rel(got, ref)
    assert abs(eq.u_max() - oeq.u_max) <= 1e-14 * oeq.u_max
    assert abs(eq.dt(0.2) - oeq.dt(0.2)) <= 1e-14 * oeq.dt(0.2)
    # k = 0 mode is zero
    assert np.all(got[(slice(None),) + (0,) * dim] == 0)


def band_limited_state(dim, n, kmax, seed):
    """Divergence-free spectrum of a real field whose modes all have |k_i| <= kmax."""
    rng = np.random.default_rng(seed)
    shape = (dim,) + (n,) * (dim - 1) + (n // 2 + 1,)
    uh = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    kf = np.abs(np.fft.fftfreq(n, 1.0 / n))
    grids = np.meshgrid(*([kf] * (dim - 1) + [np.arange(n // 2 + 1)]), indexing="ij")
    mask = np.ones(shape[1:], bool)
    for g_ in grids:
        mask &= g_ <= kmax
    uh = O.fft_r2c(O.fft_c2r(uh * mask, n)) * mask
    return O.leray(uh)


def embed(uh_small, n_small, N):
    """The same band-limited field sampled on an N-grid: identical signed modes, each
    coefficient scaled by (N / n_small)^dim (unnormalised transforms sum over N^dim points)."""
    dim = uh_small.shape[0]
    out = torch.zeros((dim,) + (N,) * (dim - 1) + (N // 2 + 1,), dtype=torch.complex128, device="cuda")
    nz = np.nonzero(np.abs(uh_small).sum(axis=0) > 0)
    idx = [np.where(a <= n_small // 2, a, a - n_small + N) for a in nz[:-1]] + [nz[-1]]
    vals = uh_small[(slice(None),) + nz] * (N / n_small) ** dim
    out[(slice(None),) + tuple(torch.from_numpy(i).cuda() for i in idx)] = dev(vals)
    return out


@pytest.mark.parametrize("N,precision,tol", [(256, "fp64", 1e-12), (512, "fp64", 1e-12),
                                             (256, "fp32_mixed", 2e-5), (512, "fp32_mixed", 3e-5)])
def test_dudt_3d_full_size_band_limited(cuda, N, precision, tol):
    """Full-size 3D RHS (every row line of the padded grid carries data) against the
    oracle on a 32^3 grid: a field with |k_i| <= 3 has products with |k_i| <= 6, which
    both grids resolve without aliasing, so without viscosity the N^3 RHS is the 32^3
    RHS on the same signed modes, scaled by (N/32)^3, and zero everywhere else.
    FP32 bar: the same RHS through float32 transforms and products on the CPU (scipy's
    pocketfft) has relative L2 error 0.6e-6 / 1.2e-6 / 2.6e-6 at N = 32 / 64 / 128, growing
    in proportion to N (the rounding floor spreads over all N^3 modes while the RHS
    sits on a few hundred), so about 1e-5 is the float32 floor at N = 512.  FP32 at
    N = 512 runs the radix-24 row kernel, at N = 256 the radix-12 one."""
    print(f"[case] 3D band-limited dudt N={N} {precision}", flush=True)
    n0 = 32
    uh0 = band_limited_state(3, n0, 3, seed=21)
    ref0 = O.IncompressibleEuler(O.Grid(3, n0), O.SmoothCutoff1D(0.0, 1.0, 1.0)).dudt(uh0)
    eq = A.IncompressibleEuler(A.Grid(3, N), A.SmoothCutoff1D(0.0, 1.0, 1.0), precision)
    u = embed(uh0, n0, N)
    out = torch.empty_like(u)
    eq.dudt(out, u)
    expect = embed(ref0, n0, N)
    torch.cuda.synchronize()
    err = float(torch.linalg.vector_norm(out - expect) / torch.linalg.vector_norm(expect))
    print(f"3D band-limited N={N} {precision}: rel L2 {err:.3e}")
    assert err <= tol


FFT_WORKERS = 16  # host threads for the oracle's 768^3 transforms


@pytest.fixture(scope="module")
def broadband_512():
    """White noise at 512^3, Leray-projected: every mode up to the Nyquist modes
    carries data."""
    print("[case] 3D broadband state N=512", flush=True)
    rng = np.random.default_rng(7)
    return O.leray(O.fft_r2c(rng.standard_normal((3, 512, 512, 512)), FFT_WORKERS))


def oracle_512():
    N = 512
    return O.IncompressibleEuler(O.Grid(3, N), O.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4),
                                 workers=FFT_WORKERS)


@pytest.fixture(scope="module")
def broadband_512_rhs(broadband_512):
    """The oracle's RHS of the broadband state and its u_max (about 20 s, 75 GB peak)."""
    print("[case] oracle RHS of the 512^3 broadband state", flush=True)
    oeq = oracle_512()
    return oeq.dudt(broadband_512), oeq.u_max


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32_mixed", 1e-5)])
def test_dudt_3d_512_broadband_against_oracle(cuda, broadband_512, broadband_512_rhs, precision, tol):
    """The benchmark's own kernels at full size (M = 768: the radix-24 row kernels, the
    768-point column passes, the k-space kernel) against the FP64 oracle on a broadband
    field, in FP64 (round-off) and FP32-mixed (float32 transforms and products; bar
    1e-5).  The oracle runs with scipy.fft's threaded pocketfft (pinned against
    brute-force DFTs like numpy's)."""
    N = 512
    print(f"[case] 3D broadband dudt N={N} {precision} against the oracle", flush=True)
    ref, u_max = broadband_512_rhs
    g, eq, _ = make_pair(3, N, precision)
    u = dev(broadband_512)
    out = torch.empty_like(u)
    eq.dudt(out, u)
    got = host(out)
    del u, out
    err = rel(got, ref)
    print(f"3D broadband N={N} {precision}: rel L2 {err:.3e}", flush=True)
    assert err <= tol, err
    assert abs(eq.u_max() - u_max) <= (1e-14 if precision == "fp64" else 1e-6) * u_max


@pytest.mark.parametrize("plan", [1, 2])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32_mixed", 1e-5)])
def test_radix_plans_3d_512_broadband_against_oracle(cuda, broadband_512, broadband_512_rhs, precision, tol,
                                                     plan):
    """Every radix plan the autotuner may select for the 768-point column passes
    (12.8.8, 4.8.24 beside the default 8.8.12), at full size against the FP64 oracle."""
    N = 512
    print(f"[case] 3D broadband dudt N={N} {precision} radix plan {plan}", flush=True)
    ref, _ = broadband_512_rhs
    g, eq, _ = make_pair(3, N, precision)
    eq.set_plans([plan, plan, 0, plan, plan, 0])
    u = dev(broadband_512)
    out = torch.empty_like(u)
    eq.dudt(out, u)
    got = host(out)
    del u, out
    err = rel(got, ref)
    print(f"3D broadband N={N} {precision} plan {plan}: rel L2 {err:.3e}", flush=True)
    assert err <= tol, err


@pytest.mark.parametrize("dim,N", [(2, 256), (2, 512), (2, 1024), (3, 256)])
@pytest.mark.parametrize("plan", [0, 1, 2])
def test_radix_plans_against_oracle(cuda, dim, N, plan):
    """Each radix plan of the 384-, 768- and 1536-point column passes against the
    oracle (FP64, 1e-12); the mixed plans per family as the autotuner may pick them."""
    print(f"[case] dudt {dim}D N={N} radix plan {plan}", flush=True)
    uh = random_state(dim, N)
    g, eq, oeq = make_pair(dim, N)
    if dim == 3:
        oeq = O.IncompressibleEuler(O.Grid(3, N), O.baseline_visc(N), workers=FFT_WORKERS)
    ref = oeq.dudt(uh)
    u = dev(uh)
    for plans in ([plan, plan, 0, plan, plan, 0], [plan, (plan + 1) % 3, 0, (plan + 2) % 3, plan, 0]):
        eq.set_plans(plans)
        out = torch.empty_like(u)
        eq.dudt(out, u)
        err = rel(host(out), ref)
        assert err <= 1e-12, (plans, err)
        assert eq.get_plans() == plans


def test_rk3_step_3d_512_broadband_against_oracle(cuda, broadband_512):
    """One SSP-RK3 step at 512^3 (three RHS, the fused stage updates, the device CFL
    step) against the oracle's step from the same broadband state."""
    N = 512
    print(f"[case] 3D broadband RK3 step N={N} fp64 against the oracle", flush=True)
    uh = broadband_512
    g, eq, _ = make_pair(3, N)
    rk = A.SSP_RK3(g.shape_fourier(3), eq)
    u = dev(uh)
    dt = rk.integrate(1.0, 1.0, 0.2, u)
    got = host(u)
    del u, rk
    print("[case] GPU step done; oracle step (three RHS)", flush=True)
    ref, dt_ref = O.ssp_rk3_integrate(oracle_512(), uh, 1.0, 0.2)
    err = rel(got, ref)
    print(f"3D broadband RK3 step N={N}: dt {dt:.6e} rel L2 {err:.3e}", flush=True)
    assert abs(dt - dt_ref) <= 1e-14 * dt_ref
    assert err <= 1e-12, err


def test_row_kernel_tiles_identical_at_full_size(cuda):
    """At M = 768 the FP64 row kernel runs with one line pair per workgroup (the
    default) or two (pdiv = 1), and the column passes with tiles of 8, 4 or 2
    columns (tsdiv 1, 2, 4): shapes the autotuner may pick.  All use the same plans
    and arithmetic per line or column, so the RHS must be bit-identical."""
    N = 512
    print(f"[case] row kernel tiles N={N}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(22)
    u = torch.randn((3, N, N, N // 2 + 1), dtype=torch.complex128, device="cuda", generator=g)
    eq = A.IncompressibleEuler(A.Grid(3, N), A.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4))
    ref = None
    for tsdiv, pdiv in (([1] * 6, 2), ([1] * 6, 1), ([2] * 6, 2), ([4] * 6, 2)):
        eq.set_tuning(tsdiv, pdiv)
        assert eq.get_tuning() == (tsdiv, pdiv)
        out = torch.empty_like(u)
        eq.dudt(out, u)
        if ref is None:
            ref = out
        else:
            torch.cuda.synchronize()
            assert torch.equal(out, ref), (tsdiv, pdiv)


def test_dudt_in_place(cuda):
    """dudt(x, x) aliasing contract (test/azeban/equations/incompressible_euler.cpp:194)."""
    uh = random_state(2, 128)
    g, eq, oeq = make_pair(2, 128)
    u = dev(uh)
    eq.dudt(u, u)
    assert rel(host(u), oeq.dudt(uh)) <= 1e-12


def test_taylor_green_2d_steady(cuda):
    """2D Taylor-Green is a steady state of inviscid Euler and its modes sit below the
    viscosity cutoff, so dudt must vanish to round-off at full size (N = 1024)."""
    N = 1024
    print("[case] TG2D steady N=1024", flush=True)
    g, eq, _ = make_pair(2, N)
    u = A.initialize(I.taylor_green_2d(N))
    out = torch.empty_like(u)
    eq.dudt(out, u)
    assert float(out.abs().max()) <= 1e-11 * float(u.abs().max())
    assert abs(eq.u_max() - (N / (1.5 * N)) ** 2) <= 1e-13


# ---------------------------------------------------------------------------
# RK3 steps and trajectories
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dim,N", [(2, 128), (3, 64)])
def test_rk3_step(cuda, dim, N):
    uh = random_state(dim, N, seed=5)
    g, eq, oeq = make_pair(dim, N)
    rk = A.SSP_RK3(g.shape_fourier(dim), eq)
    u = dev(uh)
    dt = rk.integrate(1.0, 1.0, 0.2, u)
    ref, dt_ref = O.ssp_rk3_integrate(oeq, uh, 1.0, 0.2)
    assert abs(dt - dt_ref) <= 1e-15 * dt_ref
    assert rel(host(u), ref) <= 1e-12


def test_trajectory_2d_dsl_100_steps(cuda):
    """Double shear layer at N = 128, 100 steps: FP64 relative L2 <= 1e-11."""
    N = 128
    print("[case] DSL N=128 100 steps", flush=True)
    uh = O.initialize(I.double_shear_layer(N, seed=1))
    g, eq, oeq = make_pair(2, N)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(2), eq), dev(uh))
    osim = O.Simulation(oeq, uh, 0.2)
    steps = sim.simulate_until(1.0, max_steps=100)
    osim.simulate_until(1.0, max_steps=100)
    assert steps == 100
    assert abs(sim.time() - osim.time) <= 1e-12
    assert rel(host(sim.u()), osim.u) <= 1e-11
    assert np.allclose(A.energy_spectrum(sim.u()), O.energy_spectrum(osim.u, N), rtol=1e-10, atol=1e-30)
    # the velocity stays divergence-free to round-off
    assert spectral_divergence(host(sim.u())) <= 1e-13


def test_taylor_green_2d_128_to_t1(cuda):
    """The CPU-runnable reference case: 2D Taylor-Green, N = 128, SSP-RK3 with
    C = 0.2 and SmoothCutoff(0.05/N, 1, 2 pi N/4) to t = 1 (about 285 steps), device
    loop against the FP64 oracle's host loop: same step count and time, relative L2
    <= 1e-11, the same energy spectrum, and a divergence-free velocity."""
    N = 128
    print("[case] TG2D N=128 to t=1", flush=True)
    uh = O.initialize(I.taylor_green_2d(N))
    g, eq, oeq = make_pair(2, N)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(2), eq), dev(uh))
    osim = O.Simulation(oeq, uh, 0.2)
    steps = sim.simulate_until(1.0)
    osim.simulate_until(1.0)
    assert steps == len(osim.dts) and steps > 250
    assert sim.time() == osim.time == 1.0
    got = host(sim.u())
    assert rel(got, osim.u) <= 1e-11
    # all energy sits in the shell K = 1; the other shells hold round-off only
    e_ref = O.energy_spectrum(osim.u, N)
    assert np.allclose(A.energy_spectrum(sim.u()), e_ref, rtol=1e-10, atol=1e-12 * e_ref.sum())
    assert spectral_divergence(got) <= 1e-13


def test_trajectory_3d_tg_20_steps(cuda):
    N = 64
    print("[case] TG3D N=64 20 steps", flush=True)
    uh = O.initialize(I.taylor_green_3d(N))
    g, eq, oeq = make_pair(3, N)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(3), eq), dev(uh))
    osim = O.Simulation(oeq, uh, 0.2)
    sim.simulate_until(1.0, max_steps=20)
    osim.simulate_until(1.0, max_steps=20)
    assert rel(host(sim.u()), osim.u) <= 1e-11
    assert spectral_divergence(host(sim.u())) <= 1e-13


@pytest.mark.parametrize("dim,N,steps,ic", [(3, 128, 4, "tg"), (2, 1024, 5, "dsl")])
def test_trajectory_large_grids(cuda, dim, N, steps, ic):
    """A few SSP-RK3 steps on the larger grids (3D 128^3: M = 192 kernels; 2D 1024^2:
    M = 1536 kernels with the split forward exchange) against the FP64 oracle."""
    print(f"[case] trajectory dim={dim} N={N} {ic} {steps} steps", flush=True)
    uh = O.initialize(I.taylor_green_3d(N) if ic == "tg" else I.double_shear_layer(N, seed=3))
    g, eq, oeq = make_pair(dim, N)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(dim), eq), dev(uh))
    osim = O.Simulation(oeq, uh, 0.2)
    assert sim.simulate_until(1.0, max_steps=steps) == steps
    osim.simulate_until(1.0, max_steps=steps)
    assert abs(sim.time() - osim.time) <= 1e-14
    assert rel(host(sim.u()), osim.u) <= 1e-12


def test_trajectory_dsl_1024_100_steps(cuda):
    """The 2D double-shear-layer configuration at its full size: N = 1024^2, FP64, 100
    SSP-RK3 steps of the device-driven loop against the oracle (threaded transforms):
    relative L2 <= 1e-11 in velocity, the same time and energy spectrum, divergence at
    round-off."""
    N, steps = 1024, 100
    print(f"[case] DSL N={N} {steps} steps against the oracle", flush=True)
    uh = O.initialize(I.double_shear_layer(N, seed=1))
    g, eq, _ = make_pair(2, N)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(2), eq), dev(uh))
    assert sim.simulate_until(1.0, max_steps=steps) == steps
    got = host(sim.u())
    oeq = O.IncompressibleEuler(O.Grid(2, N), O.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4),
                                workers=FFT_WORKERS)
    osim = O.Simulation(oeq, uh, 0.2)
    for done in range(20, steps + 1, 20):
        osim.simulate_until(1.0, max_steps=20)
        print(f"oracle: {done} steps, t = {osim.time:.6e}", flush=True)
    assert len(osim.dts) == steps
    assert abs(sim.time() - osim.time) <= 1e-12 * osim.time
    err = rel(got, osim.u)
    print(f"DSL N={N} after {steps} steps: rel L2 {err:.3e}", flush=True)
    assert err <= 1e-11, err
    ek = np.asarray(A.energy_spectrum(sim.u()))
    ek_ref = O.energy_spectrum(osim.u, N)
    assert np.allclose(ek, ek_ref, rtol=1e-10, atol=1e-12 * float(np.sum(ek_ref)))
    assert spectral_divergence(got) <= 1e-13


@pytest.mark.timeout(400)  # about two minutes, most of it the oracle's 300 RHS
def test_trajectory_3d_fbm_128_100_steps(cuda):
    """A broadband 3D state (fractional Brownian, H = 0.75) at 128^3, 100 SSP-RK3 steps
    of the device-driven loop against the oracle (threaded transforms): relative L2
    <= 1e-11 in velocity, the same time and energy spectrum, divergence at round-off."""
    N, steps = 128, 100
    uh = random_state(3, N, seed=9)
    print(f"[case] 3D fBm N={N} {steps} steps against the oracle", flush=True)
    g, eq, _ = make_pair(3, N)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(3), eq), dev(uh))
    assert sim.simulate_until(1.0, max_steps=steps) == steps
    got = host(sim.u())
    oeq = O.IncompressibleEuler(O.Grid(3, N), O.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4),
                                workers=FFT_WORKERS)
    osim = O.Simulation(oeq, uh, 0.2)
    for done in range(20, steps + 1, 20):
        osim.simulate_until(1.0, max_steps=20)
        print(f"oracle: {done} steps, t = {osim.time:.6e}", flush=True)
    assert len(osim.dts) == steps
    assert abs(sim.time() - osim.time) <= 1e-12 * osim.time
    err = rel(got, osim.u)
    print(f"3D fBm N={N} after {steps} steps: rel L2 {err:.3e}", flush=True)
    assert err <= 1e-11, err
    ek = np.asarray(A.energy_spectrum(sim.u()))
    ek_ref = O.energy_spectrum(osim.u, N)
    assert np.allclose(ek, ek_ref, rtol=1e-10, atol=1e-12 * float(np.sum(ek_ref)))
    assert spectral_divergence(got) <= 1e-13


def test_simulate_until_device_loop_matches_host_loop(cuda):
    """The device-driven loop (dt and time on the device) takes the same steps as the
    host loop of Simulation::simulate_until, including the final clamp onto t."""
    N = 64
    uh = O.initialize(I.double_shear_layer(N, seed=2))
    g, eq, _ = make_pair(2, N)
    a = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(2), eq), dev(uh))
    g2, eq2, _ = make_pair(2, N)
    b = A.Simulation(g2, 0.2, A.SSP_RK3(g2.shape_fourier(2), eq2), dev(uh))
    na = a.simulate_until(0.05, chunk=4)
    nb = b.simulate_until_host(0.05)
    assert na == nb
    assert a.time() == b.time() == 0.05
    assert rel(host(a.u()), host(b.u())) <= 1e-14


def test_simulate_until_edge_cases(cuda):
    """Simulation::simulate_until (src/azeban/simulation.cpp:89-102): no step when the
    target is not ahead, time_ = t afterwards even if t lies behind, a target closer
    than one CFL step is reached in one clamped step, and max_steps bounds the loop."""
    N = 64
    uh = O.initialize(I.double_shear_layer(N, seed=4))
    g, eq, _ = make_pair(2, N)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(2), eq), dev(uh))
    u0 = host(sim.u()).copy()
    assert sim.simulate_until(0.0) == 0 and sim.time() == 0.0
    assert np.array_equal(host(sim.u()), u0)
    sim.set_time(0.5)
    assert sim.simulate_until(0.25) == 0 and sim.time() == 0.25
    assert np.array_equal(host(sim.u()), u0)
    sim.reset()
    t_small = 1e-5  # far below one CFL step at N = 64
    g2, eq2, _ = make_pair(2, N)
    ref = A.Simulation(g2, 0.2, A.SSP_RK3(g2.shape_fourier(2), eq2), dev(uh))
    assert sim.simulate_until(t_small) == 1 and ref.simulate_until_host(t_small) == 1
    assert sim.time() == t_small == ref.time()
    assert rel(host(sim.u()), host(ref.u())) <= 1e-15
    n = sim.simulate_until(1.0, max_steps=3, chunk=2)
    assert n == 3 and t_small < sim.time() < 1.0


# every FP32 row-kernel path: 2D L = 384 (row-2 kernel), 768 (two line pairs per
# workgroup), 1536 (24 values per thread), 3D L = 96 and 192 (row-4 kernel), 384
# (single forward round)
@pytest.mark.parametrize("dim,N,steps", [(2, 256, 20), (2, 512, 10), (2, 1024, 10), (3, 64, 10), (3, 128, 5),
                                         (3, 256, 2)])
def test_fp32_mixed_error(cuda, dim, N, steps):
    print(f"[case] fp32_mixed dim={dim} N={N}", flush=True)
    uh = random_state(dim, N, seed=6)
    g, eq, oeq = make_pair(dim, N, "fp32_mixed")
    if dim == 3 and N >= 256:
        oeq = O.IncompressibleEuler(O.Grid(3, N), O.baseline_visc(N), workers=FFT_WORKERS)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(dim), eq), dev(uh))
    osim = O.Simulation(oeq, uh, 0.2)
    sim.simulate_until(10.0, max_steps=steps)
    osim.simulate_until(10.0, max_steps=steps)
    err = rel(host(sim.u()), osim.u)
    print(f"fp32_mixed rel L2 vs FP64 oracle after {steps} steps: {err:.3e}")
    assert err <= 1e-5


def test_fp32_mixed_taylor_green_512_100_steps(cuda):
    """The benchmark's FP32-mixed form at its own size: 3D Taylor-Green 512^3,
    100 SSP-RK3 steps, against the FP64 path (which the smaller-size tests tie to
    the oracle at 1e-11): relative L2 <= 1e-5, and the same step count."""
    N = 512
    print(f"[case] fp32_mixed TG3D N={N} 100 steps", flush=True)
    u0 = A.initialize(I.taylor_green_3d(N))
    res = {}
    for prec in ("fp64", "fp32_mixed"):
        g, eq, _ = make_pair(3, N, prec)
        sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(3), eq), u0)
        n = sim.simulate_until(10.0, max_steps=100, chunk=20)
        assert n == 100
        res[prec] = (sim.u().clone(), sim.time())
        del sim, eq
        torch.cuda.empty_cache()
    a, ta = res["fp64"]
    b, tb = res["fp32_mixed"]
    err = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(a))
    print(f"fp32_mixed rel L2 vs FP64 after 100 steps at 512^3: {err:.3e}; t {ta} vs {tb}")
    assert err <= 1e-5
    assert abs(ta - tb) <= 1e-6 * ta


def test_copy_bandwidth_probe(cuda):
    """The streaming-copy reference used for the measured roofline returns a sane rate."""
    import ctypes as C

    from azeban_amd._capi import check, load

    out = C.c_double()
    check(load().azb_copy_bandwidth(1.0, 3, C.byref(out)))
    print(f"stream copy: {out.value:.0f} GB/s")
    assert 500.0 < out.value < 20000.0
