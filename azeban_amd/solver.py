"""Python mirror of azeban's Equation / TimeIntegrator / Simulation API for the
IncompressibleEuler<Dim, SmoothCutoff1D, NoForcing> + SSP_RK3 path, running on
the HIP library through its C ABI.  Device arrays are torch tensors (complex128
on the GPU); torch is used for device memory only.

Reference interfaces mirrored:
  Grid<Dim>                         include/azeban/grid.hpp:30-53
  SmoothCutoff1D / factory          include/azeban/equations/spectral_viscosity.hpp:56-77,
                                    include/azeban/equations/spectral_viscosity_factory.hpp:29-54
  IncompressibleEuler::dudt, dt     include/azeban/equations/incompressible_euler.hpp:103-124
  SSP_RK3::integrate                include/azeban/evolution/ssp_rk3.hpp:48-64
  Simulation::simulate_until        src/azeban/simulation.cpp:89-102
  FFT forward/backward              include/azeban/operations/fft.hpp:37-121
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import _capi
from ._capi import AZB_FP32_MIXED, AZB_FP64, SVParams, check, load

PRECISIONS = {"fp64": AZB_FP64, "fp32_mixed": AZB_FP32_MIXED}


def _ptr(t: torch.Tensor) -> C.c_void_p:
    if not t.is_cuda:
        raise ValueError("expected a device (cuda/hip) tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _need(name: str, t: torch.Tensor, dtypes, shape_ok=None, what: str = "") -> None:
    """Reject a tensor of the wrong dtype or shape before any library call: the entry
    points take a bare pointer and an element count, so a mismatch would read or write
    outside the tensor on the device."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor")
    if t.dtype not in dtypes:
        raise ValueError(f"{name}: expected dtype {' or '.join(map(str, dtypes))}, got {t.dtype}")
    if shape_ok is not None and not shape_ok(tuple(t.shape)):
        raise ValueError(f"{name}: expected shape {what}, got {tuple(t.shape)}")


def _fourier_shape_ok(shape) -> bool:
    """(dim, N, ..., N/2+1) with dim 2 or 3 and an even N."""
    if len(shape) not in (3, 4) or shape[0] != len(shape) - 1:
        return False
    N = shape[1]
    return N >= 2 and N % 2 == 0 and all(n == N for n in shape[1:-1]) and shape[-1] == N // 2 + 1


def _component_shape_ok(shape) -> bool:
    """One component (N, ..., N/2+1) in 2 or 3 dimensions with an even N."""
    if len(shape) not in (2, 3):
        return False
    N = shape[0]
    return N >= 2 and N % 2 == 0 and all(n == N for n in shape[:-1]) and shape[-1] == N // 2 + 1


_FOURIER = "(dim, N, ..., N/2+1) with dim = number of axes after the first"
_COMPONENT = "(N, ..., N/2+1) in 2 or 3 dimensions"


@dataclass
class Grid:
    """Grid<Dim>: N_phys, N_phys_pad (default 3N/2), N_fourier = N/2+1."""

    dim: int
    N_phys: int
    N_phys_pad: int = 0

    def __post_init__(self):
        if self.N_phys_pad == 0:
            self.N_phys_pad = 3 * self.N_phys // 2
        self.N_fourier = self.N_phys // 2 + 1
        self.N_fourier_pad = self.N_phys_pad // 2 + 1

    def shape_fourier(self, n_vars: int):
        return (n_vars,) + (self.N_phys,) * (self.dim - 1) + (self.N_fourier,)

    def shape_phys(self, n_vars: int):
        return (n_vars,) + (self.N_phys,) * self.dim


@dataclass
class SmoothCutoff1D:
    eps: float
    s: float
    k0: float


def make_smooth_cutoff_1d(config: dict, grid: Grid) -> SmoothCutoff1D:
    """Factory of spectral_viscosity_factory.hpp:29-54 (missing "eps" is an error)."""
    s = float(config.get("s", 1.0))
    theta = float(config.get("theta", (2 * s - 1) / (2 * s)))
    m0 = float(config.get("m0", 2 * math.pi))
    if "eps" not in config:
        raise ValueError('Smooth Cutoff Viscosity expects "eps" parameter in config')
    eps = float(config["eps"])
    return SmoothCutoff1D(eps / float(grid.N_phys) ** (2 * s - 1), s, m0 * float(grid.N_phys) ** theta)


class IncompressibleEuler:
    """Equation<Dim> for incompressible Euler with SmoothCutoff1D viscosity, no forcing.

    Owns the HIP solver handle (workspace, stage buffers, FFT twiddles).
    """

    def __init__(self, grid: Grid, visc: SmoothCutoff1D, precision: str = "fp64"):
        self.grid = grid
        self.visc_ = visc
        self.precision = precision
        lib = load()
        h = C.c_void_p()
        check(lib.azb_create(grid.dim, grid.N_phys, grid.N_phys_pad, grid.dim, PRECISIONS[precision],
                             SVParams(visc.eps, visc.s, visc.k0), _stream(), C.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                load().azb_destroy(h)
            except Exception:  # interpreter teardown: the process releases the device anyway
                pass
            self.handle = None

    def n_vars(self) -> int:
        return self.grid.dim

    def visc(self) -> float:
        return self.visc_.eps

    def workspace_bytes(self) -> int:
        return int(load().azb_workspace_bytes(self.handle))

    def dudt(self, dudt_hat: torch.Tensor, u_hat: torch.Tensor, t: float = 0.0, dt: float = 0.0,
             C_: float = 0.0) -> None:
        """dudt_hat <- L(u_hat); dudt_hat may be u_hat itself."""
        self._check(u_hat)
        self._check(dudt_hat)
        check(load().azb_dudt(self.handle, _ptr(u_hat), _ptr(dudt_hat)))

    def dt(self, C_: float) -> float:
        out = C.c_double()
        check(load().azb_dt(self.handle, C_, C.byref(out)))
        return out.value

    def u_max(self) -> float:
        out = C.c_double()
        check(load().azb_u_max(self.handle, C.byref(out)))
        return out.value

    # launch-shape tuning: column tile divisors per kernel family, row line-pair divisor
    def get_tuning(self):
        ts = (C.c_int * len(_capi.PROF_FAMILIES))()
        pd = C.c_int()
        check(load().azb_get_tuning(self.handle, ts, C.byref(pd)))
        return list(ts), pd.value

    def set_tuning(self, tsdiv, pdiv: int) -> None:
        ts = (C.c_int * len(_capi.PROF_FAMILIES))(*tsdiv)
        check(load().azb_set_tuning(self.handle, ts, pdiv))

    # radix plans of the column passes (results change at round-off with the plan)
    def get_plans(self):
        pl = (C.c_int * len(_capi.PROF_FAMILIES))()
        check(load().azb_get_plans(self.handle, pl))
        return list(pl)

    def set_plans(self, plans) -> None:
        check(load().azb_set_plans(self.handle, (C.c_int * len(_capi.PROF_FAMILIES))(*plans)))

    def autotune(self, u_hat: torch.Tensor, reps: int = 3, cache: str | None = None, nsamples: int = 1):
        """Pick the radix plans, then the launch shapes, by timing on the device.  The plans
        are chosen once per (dim, N, precision) and shared by every solver of that grid
        in the process (and in `cache`), so that batched and single solves of one grid
        keep computing bit-identical results; the launch shapes are chosen per
        (dim, N, precision, samples)."""
        _ensure_plans(self.handle, self.grid.dim, self.grid.N_phys, self.precision, u_hat, reps, cache)
        key = f"{self.grid.dim}/{self.grid.N_phys}/{self.precision}/{nsamples}"
        store = _tuning_load(cache) if cache else {}
        plans = self.get_plans()
        if key in store:
            self.set_tuning(store[key]["tsdiv"], store[key]["pdiv"])
            return dict(store[key], plans=plans)
        check(load().azb_autotune(self.handle, _ptr(u_hat), reps))
        ts, pd = self.get_tuning()
        rec = {"tsdiv": ts, "pdiv": pd}
        if cache:
            store = _tuning_load(cache)
            store[key] = rec
            _tuning_save(cache, store)
        return dict(rec, plans=plans)

    # per-kernel-family hipEvent timing on the solver's stream
    def profile_enable(self, on: bool = True) -> None:
        check(load().azb_profile_enable(self.handle, 1 if on else 0))

    def profile_reset(self) -> None:
        check(load().azb_profile_reset(self.handle))

    def profile_read(self) -> dict:
        n = len(_capi.PROF_FAMILIES)
        ms = (C.c_double * n)()
        cnt = (C.c_long * n)()
        check(load().azb_profile_read(self.handle, ms, cnt, n))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_capi.PROF_FAMILIES)}

    def _check(self, u: torch.Tensor):
        if tuple(u.shape) != self.grid.shape_fourier(self.n_vars()) or u.dtype != torch.complex128:
            raise ValueError(f"expected complex128 of shape {self.grid.shape_fourier(self.n_vars())}, "
                             f"got {u.dtype} {tuple(u.shape)}")


# radix plans per (dim, N, precision), shared by every solver of the process
_PLANS: dict = {}


def _ensure_plans(handle, dim: int, N: int, precision: str, u_hat: torch.Tensor, reps: int,
                  cache: str | None) -> list:
    """Give the solver the plans of its grid: from this process, from `cache`, or timed
    now on the device (azb_autotune_plans) and remembered."""
    lib = load()
    key = f"{dim}/{N}/{precision}"
    plans = _PLANS.get(key)
    if plans is None and cache:
        plans = _tuning_load(cache).get("plans", {}).get(key)
    n = len(_capi.PROF_FAMILIES)
    if plans is None:
        check(lib.azb_autotune_plans(handle, C.c_void_p(u_hat.data_ptr()), reps))
        pl = (C.c_int * n)()
        check(lib.azb_get_plans(handle, pl))
        plans = list(pl)
        if cache:
            store = _tuning_load(cache)
            store.setdefault("plans", {})[key] = plans
            _tuning_save(cache, store)
    else:
        check(lib.azb_set_plans(handle, (C.c_int * n)(*plans)))
    _PLANS[key] = list(plans)
    return list(plans)


def _tuning_load(path: str) -> dict:
    import json
    import os

    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _tuning_save(path: str, store: dict) -> None:
    import json
    import os

    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(store, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


class SSP_RK3:
    """TimeIntegrator<Dim>: Shu-Osher SSP-RK3 with dt from the first stage's CFL."""

    def __init__(self, shape, equation: IncompressibleEuler):
        self.shape = tuple(shape)
        self.equation_ = equation

    def equation(self) -> IncompressibleEuler:
        return self.equation_

    def integrate(self, t: float, max_dt: float, C_: float, u: torch.Tensor) -> float:
        self.equation_._check(u)
        out = C.c_double()
        check(load().azb_rk3_step(self.equation_.handle, _ptr(u), max_dt, C_, C.byref(out)))
        return out.value


class Simulation:
    """Simulation<Dim>: owns u_hat, runs the time loop to a target time."""

    def __init__(self, grid: Grid, C_: float, timestepper: SSP_RK3, u_hat: torch.Tensor | None = None):
        self.grid = grid
        self.C = C_
        self.timestepper = timestepper
        n = timestepper.equation().n_vars()
        if u_hat is None:
            self.u_ = torch.zeros(grid.shape_fourier(n), dtype=torch.complex128, device="cuda")
        else:
            self.u_ = u_hat.detach().to(device="cuda", dtype=torch.complex128).clone().contiguous()
        self.time_ = 0.0
        self.steps = 0

    def u(self) -> torch.Tensor:
        return self.u_

    def time(self) -> float:
        return self.time_

    def set_time(self, t: float):
        self.time_ = t

    def reset(self):
        self.time_ = 0.0

    def simulate_until(self, t: float, max_steps: int = 0, chunk: int = 8) -> int:
        """Device-driven loop (dt stays on the device); returns the number of steps."""
        eq = self.timestepper.equation()
        time = C.c_double(self.time_)
        steps = C.c_int(0)
        check(load().azb_simulate_until(eq.handle, _ptr(self.u_), t, self.C, C.byref(time), max_steps,
                                        chunk, C.byref(steps)))
        self.time_ = time.value
        self.steps += steps.value
        return steps.value

    def simulate_for(self, t: float) -> int:
        return self.simulate_until(self.time_ + t)

    def simulate_until_host(self, t: float) -> int:
        """Host-driven loop, step by step exactly as Simulation::simulate_until does."""
        eq = self.timestepper.equation()
        n = 0
        while self.time_ < t:
            max_dt = min(t - self.time_, self.C * 2.0 / (eq.visc() * (math.pi * self.grid.N_phys) ** 2))
            dt = self.timestepper.integrate(t, max_dt, self.C, self.u_)
            self.time_ += dt
            n += 1
        self.time_ = t
        self.steps += n
        return n


# ---------------------------------------------------------------------------
# functional seams
# ---------------------------------------------------------------------------
def fft_r2c(u: torch.Tensor) -> torch.Tensor:
    """Unnormalised R2C over all but the leading (batch) axis."""
    _need("fft_r2c", u, (torch.float64, torch.float32),
          lambda sh: 2 <= len(sh) <= 4 and sh[0] >= 1 and len(set(sh[1:])) == 1, "(howmany, n, ..., n)")
    dim = u.dim() - 1
    n = u.shape[-1]
    prec = AZB_FP64 if u.dtype == torch.float64 else AZB_FP32_MIXED
    cdt = torch.complex128 if u.dtype == torch.float64 else torch.complex64
    out = torch.empty(tuple(u.shape[:-1]) + (n // 2 + 1,), dtype=cdt, device=u.device)
    u = u.contiguous()
    check(load().azb_fft_r2c(dim, n, u.shape[0], prec, _ptr(u), _ptr(out), _stream()))
    return out


def fft_c2r(u_hat: torch.Tensor, n: int) -> torch.Tensor:
    """Unnormalised C2R; works on a copy (the library overwrites its input)."""
    _need("fft_c2r", u_hat, (torch.complex128, torch.complex64),
          lambda sh: 2 <= len(sh) <= 4 and sh[0] >= 1 and all(x == n for x in sh[1:-1]) and sh[-1] == n // 2 + 1,
          f"(howmany, {n}, ..., {n // 2 + 1})")
    dim = u_hat.dim() - 1
    prec = AZB_FP64 if u_hat.dtype == torch.complex128 else AZB_FP32_MIXED
    rdt = torch.float64 if u_hat.dtype == torch.complex128 else torch.float32
    tmp = u_hat.contiguous().clone()
    out = torch.empty(tuple(u_hat.shape[:-1]) + (n,), dtype=rdt, device=u_hat.device)
    check(load().azb_fft_c2r(dim, n, u_hat.shape[0], prec, _ptr(tmp), _ptr(out), _stream()))
    return out


def leray_(u_hat: torch.Tensor) -> torch.Tensor:
    _need("leray_", u_hat, (torch.complex128,), _fourier_shape_ok, _FOURIER)
    dim = u_hat.dim() - 1
    check(load().azb_leray(dim, u_hat.shape[1], _ptr(u_hat), _stream()))
    return u_hat


def initialize(u_phys) -> torch.Tensor:
    """Initializer::initialize: u_hat = FFT(u) (unnormalised) then Leray, on the device."""
    u = torch.as_tensor(u_phys, dtype=torch.float64).to("cuda").contiguous()
    return leray_(fft_r2c(u))


def energy_spectrum(u_hat: torch.Tensor):
    _need("energy_spectrum", u_hat, (torch.complex128,), _fourier_shape_ok, _FOURIER)
    dim = u_hat.dim() - 1
    N = u_hat.shape[1]
    spec = (C.c_double * (N // 2 + 1))()
    check(load().azb_energy_spectrum(dim, N, _ptr(u_hat), spec, _stream()))
    return list(spec)


def copy_to_padded(src: torch.Tensor, M: int) -> torch.Tensor:
    """One component, N -> M zero padding."""
    _need("copy_to_padded", src, (torch.complex128,), _component_shape_ok, _COMPONENT)
    if M < src.shape[0]:
        raise ValueError("copy_to_padded: M must not be smaller than N")
    dim = src.dim()
    N = src.shape[0]
    shape = (M,) * (dim - 1) + (M // 2 + 1,)
    out = torch.empty(shape, dtype=torch.complex128, device=src.device)
    src = src.contiguous()
    check(load().azb_copy_to_padded(dim, N, M, _ptr(src), _ptr(out), _stream()))
    return out


def copy_from_padded(src: torch.Tensor, N: int) -> torch.Tensor:
    """One component, M -> N truncation (the inverse of copy_to_padded)."""
    _need("copy_from_padded", src, (torch.complex128,), _component_shape_ok, _COMPONENT)
    if not 2 <= N <= src.shape[0] or N % 2:
        raise ValueError("copy_from_padded: N must be even and at most the padded size")
    dim = src.dim()
    M = src.shape[0]
    shape = (N,) * (dim - 1) + (N // 2 + 1,)
    out = torch.empty(shape, dtype=torch.complex128, device=src.device)
    src = src.contiguous()
    check(load().azb_copy_from_padded(dim, M, N, _ptr(src), _ptr(out), _stream()))
    return out


def scale_(a: float, x: torch.Tensor) -> torch.Tensor:
    """x <- a x in place over a float64 or complex128 tensor."""
    if not x.is_contiguous() or x.dtype not in (torch.float64, torch.complex128):
        raise ValueError("scale_ needs a contiguous float64 or complex128 tensor")
    n = x.numel() * (2 if x.is_complex() else 1)
    check(load().azb_scale(a, _ptr(x), n, _stream()))
    return x


def to_physical(u_hat: torch.Tensor, N_out: int | None = None) -> torch.Tensor:
    """Velocity in physical space as azeban writes it (run_from_config.cpp:93-98): u_hat
    scaled by 1/N^d, then the unnormalised inverse transform.  With N_out < N the
    spectrum is first truncated to N_out modes per axis, as the sample writer's
    downsampled output does (netcdf_sample_writer.cpp:143-163); the scale stays 1/N^d."""
    dim = u_hat.dim() - 1
    N = u_hat.shape[1]
    n = N if N_out is None else N_out
    if n > N:
        raise ValueError("N_out must not exceed the grid size")
    if n == N:
        tmp = u_hat.contiguous().clone()
    else:
        tmp = torch.stack([copy_from_padded(u_hat[c], n) for c in range(u_hat.shape[0])])
    scale_(1.0 / float(N) ** dim, tmp)
    return fft_c2r(tmp, n)


def axpby_(a: float, x: torch.Tensor, b: float, y: torch.Tensor) -> torch.Tensor:
    """y <- a x + b y over complex128 tensors of equal size."""
    _need("axpby_ x", x, (torch.complex128,))
    _need("axpby_ y", y, (torch.complex128,))
    if x.numel() != y.numel():
        raise ValueError(f"axpby_: x has {x.numel()} elements, y {y.numel()}")
    check(load().azb_axpby(a, _ptr(x), b, _ptr(y), x.numel(), _stream()))
    return y


def max_norm(x: torch.Tensor) -> float:
    """max |x| over a float64 tensor (the reference's max_norm of a real array)."""
    _need("max_norm", x, (torch.float64,))
    out = C.c_double()
    x = x.contiguous()
    check(load().azb_max_norm(_ptr(x), x.numel(), C.byref(out), _stream()))
    return out.value
