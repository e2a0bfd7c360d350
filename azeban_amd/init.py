"""Initial data of the benchmark workloads, as physical-space velocity fields on
the N-grid x_i = i/N of [0,1]^d (axis 0 = x).  They are turned into the Fourier
state by initialize() (device FFT + Leray), the convention of azeban's
Initializer::initialize (include/azeban/init/initializer.hpp:68-75).

Random draws use numpy's MT19937 bit generator with the given seed.
"""
from __future__ import annotations

import math

import numpy as np

TWOPI = 2.0 * math.pi


def _coords(N: int, dim: int):
    x = np.arange(N) / N
    return np.meshgrid(*([x] * dim), indexing="ij")


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.MT19937(seed))


def taylor_green_2d(N: int) -> np.ndarray:
    """u = sin(2 pi x) cos(2 pi y), v = -cos(2 pi x) sin(2 pi y)."""
    X, Y = _coords(N, 2)
    return np.stack([np.sin(TWOPI * X) * np.cos(TWOPI * Y), -np.cos(TWOPI * X) * np.sin(TWOPI * Y)])


def taylor_green_3d(N: int) -> np.ndarray:
    """u = (sin x cos y cos z, -cos x sin y cos z, 0) with x = 2 pi i / N."""
    X, Y, Z = (TWOPI * a for a in _coords(N, 3))
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    return np.stack([u, v, np.zeros_like(u)])


def double_shear_layer(N: int, seed: int = 1, rho: float = 0.05, delta: float = 0.05,
                       n_modes: int = 10, amp: float = 0.025) -> np.ndarray:
    """u = tanh((y - 1/4)/rho) for y <= 1/2 else tanh((3/4 - y)/rho);
    v = delta sin(2 pi x) + sum_{m=1..n_modes} a_m sin(2 pi m x + phi_m),
    a_m ~ U[-amp, amp], phi_m ~ U[0, 2 pi), drawn as (a_m, phi_m) per m."""
    X, Y = _coords(N, 2)
    u = np.where(Y <= 0.5, np.tanh((Y - 0.25) / rho), np.tanh((0.75 - Y) / rho))
    v = delta * np.sin(TWOPI * X)
    rng = _rng(seed)
    for m in range(1, n_modes + 1):
        a = rng.uniform(-amp, amp)
        phi = rng.uniform(0.0, TWOPI)
        v = v + a * np.sin(TWOPI * m * X + phi)
    return np.stack([u, v])


def vortex_patch(N: int, seed: int = 1, radius: float = 0.2, n_modes: int = 16,
                 amp: float = 0.01) -> np.ndarray:
    """Discontinuous vortex patch: solid-body rotation (-(y-1/2)/2, (x-1/2)/2) inside
    r < radius (1 + amp sum_m alpha_m cos(m theta + phi_m)), zero outside;
    alpha_m ~ U[-1, 1], phi_m ~ U[0, 2 pi)."""
    X, Y = _coords(N, 2)
    dx, dy = X - 0.5, Y - 0.5
    r = np.sqrt(dx * dx + dy * dy)
    theta = np.arctan2(dy, dx)
    rng = _rng(seed)
    pert = np.zeros_like(r)
    for m in range(1, n_modes + 1):
        alpha = rng.uniform(-1.0, 1.0)
        phi = rng.uniform(0.0, TWOPI)
        pert += alpha * np.cos(m * theta + phi)
    inside = r < radius * (1.0 + amp * pert)
    return np.stack([np.where(inside, -0.5 * dy, 0.0), np.where(inside, 0.5 * dx, 0.0)])


def fractional_brownian(dim: int, N: int, seed: int = 1, hurst: float = 0.75) -> np.ndarray:
    """Random field with |c_k| ~ |k|^-(H + d/2) and Gaussian coefficients per mode and
    component; real by construction (inverse real FFT of a half spectrum), zero mean,
    scaled to unit RMS speed.  Divergence is removed later by the Leray projection."""
    rng = _rng(seed)
    shape = (N,) * (dim - 1) + (N // 2 + 1,)
    kf = np.fft.fftfreq(N, 1.0 / N)
    kl = np.arange(N // 2 + 1)
    grids = np.meshgrid(*([kf] * (dim - 1) + [kl]), indexing="ij")
    k2 = sum(g.astype(float) ** 2 for g in grids)
    with np.errstate(divide="ignore"):
        amp = np.where(k2 > 0, k2 ** (-(hurst + dim / 2.0) / 2.0), 0.0)
    comps = []
    for _ in range(dim):
        c = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * amp
        comps.append(np.fft.irfftn(c, s=(N,) * dim, axes=tuple(range(dim))))
    u = np.stack(comps)
    u -= u.mean(axis=tuple(range(1, dim + 1)), keepdims=True)
    rms = math.sqrt(float(np.mean(np.sum(u * u, axis=0))))
    return u / rms if rms > 0 else u
