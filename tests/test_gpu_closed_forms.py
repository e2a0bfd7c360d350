"""Whole-path checks on the GPU that need no oracle run: closed forms and convergence.

* A parallel shear flow u = (sin 2pi(m y + p z), 0, 0) is a steady state of the
  inviscid Euler equations: B = u (x) u has modes only at k_1 = 0, so ik . B_hat
  vanishes and dudt = nu(|k|) u_hat exactly.  Through SSP-RK3 each step multiplies
  the state by the stability polynomial R(z) = 1 + z + z^2/2 + z^3/6, z = nu dt,
  with dt = C / (N u_max) and u_max = (N/M)^d times the amplitude times the
  largest |sin| on the padded grid.  This is the tracer-free form of the reference's spectral-viscosity
  heat test (test/azeban/equations/spectral_viscosity.cpp:49-104, 155-297), run
  through the device RHS, the device CFL step and the fused RK stages.
* Convergence in the style of test/azeban/equations/incompressible_euler.cpp:44-93:
  the double shear layer at N = 64, 128, 256 against N = 512 at t = 0.1, error in
  the L2 norm of the velocity (Parseval over the spectral coefficients, coarse
  spectra embedded by signed wavenumber), observed rate >= 1 as the reference asks.
  The FP64 oracle gives rates 4.0 and 4.4 for the same setup at N = 32..256.
"""
import math

import numpy as np
import pytest
import torch

from oracle import azeban_oracle as O

import azeban_amd as A
from azeban_amd import init as I

pytestmark = pytest.mark.gpu


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def sv(kabs, eps, s, k0):
    """SmoothCutoff1D (include/azeban/equations/spectral_viscosity.hpp:27-38)."""
    return -eps * kabs ** (2 * s) * (1.0 - math.exp(-((kabs / k0) ** 18)))


def rk3_poly(z):
    return 1.0 + z + z * z / 2.0 + z ** 3 / 6.0


@pytest.mark.parametrize("dim,N,mode", [(2, 128, (40,)), (2, 128, (30,)), (2, 1024, (300,)),
                                        (3, 64, (12, 16))])
def test_shear_mode_viscous_decay(cuda, dim, N, mode):
    print(f"[case] shear mode dim={dim} N={N} mode={mode}", flush=True)
    M = 3 * N // 2
    eps, s, k0, C = 0.05 / N, 1.0, 2 * math.pi * N / 4, 0.2
    x = np.arange(N) / N
    if dim == 2:
        (m,) = mode
        X, Y = np.meshgrid(x, x, indexing="ij")
        u = np.stack([np.sin(2 * math.pi * m * Y), np.zeros_like(Y)])
        kabs = 2 * math.pi * m
        jm = np.arange(M)
        smax = float(np.abs(np.sin(2 * math.pi * m * jm / M)).max())
    else:
        m, p = mode
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        u = np.stack([np.sin(2 * math.pi * (m * Y + p * Z)), np.zeros_like(Y), np.zeros_like(Y)])
        kabs = 2 * math.pi * math.hypot(m, p)
        jm = np.arange(M)
        smax = float(np.abs(np.sin(2 * math.pi * (m * jm[:, None] + p * jm[None, :]) / M)).max())
    nu = sv(kabs, eps, s, k0)
    assert nu < -1e-3  # the mode sits where the viscosity acts

    g = A.Grid(dim, N)
    eq = A.IncompressibleEuler(g, A.SmoothCutoff1D(eps, s, k0))
    rk = A.SSP_RK3(g.shape_fourier(dim), eq)
    uh = A.initialize(u)
    u0 = host(uh).copy()
    amp = 1.0
    dt0 = None
    for n in range(12):
        # six CFL-limited steps, then six steps clamped by max_dt (ssp_rk3.hpp:48-64)
        max_dt = 1.0 if n < 6 else 0.5 * dt0
        dt = rk.integrate(1.0, max_dt, C, uh)
        # u_max = max|u| / M^d over the padded grid of the unnormalised transform
        dt_expected = min(C / (N * (N / M) ** dim * amp * smax), max_dt)
        assert abs(dt - dt_expected) <= 1e-13 * dt_expected, (n, dt, dt_expected)
        dt0 = dt0 or dt
        amp *= rk3_poly(nu * dt)
    got = host(uh)
    assert amp < 0.95  # the decay is visible
    err = np.linalg.norm((got - amp * u0).ravel()) / np.linalg.norm((amp * u0).ravel())
    assert err <= 1e-12, err


def _l2_error(uc, Nc, ur, Nr):
    """L2 norm of u_coarse - u_ref on the unit torus from the spectral coefficients:
    the coarse spectrum is embedded in the reference grid by signed wavenumber and the
    last (half) axis counts its interior modes twice (Hermitian symmetry)."""
    emb = O.copy_to_padded(uc / Nc ** 2, ur.shape)
    d = emb - ur / Nr ** 2
    w = np.full(d.shape[-1], 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    return math.sqrt(float(np.sum(w * np.abs(d) ** 2)))


def test_convergence_double_shear_layer(cuda):
    t_end, C = 0.1, 0.2
    sols = {}
    for N in (64, 128, 256, 512):
        print(f"[case] DSL convergence N={N}", flush=True)
        g = A.Grid(2, N)
        eq = A.IncompressibleEuler(g, A.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4))
        sim = A.Simulation(g, C, A.SSP_RK3(g.shape_fourier(2), eq), A.initialize(I.double_shear_layer(N, seed=1)))
        sim.simulate_until(t_end)
        assert sim.time() == t_end
        sols[N] = host(sim.u())
    errs = [_l2_error(sols[N], N, sols[512], 512) for N in (64, 128, 256)]
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print(f"errors {errs} rates {rates}", flush=True)
    assert all(r >= 1.0 for r in rates), rates
