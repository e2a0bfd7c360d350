"""Continuous-batching solve server.

Requests (a 2D or 3D state of any built size plus a target time) arrive over
time.  The server admits a waiting request as soon as its device footprint
(azb_bytes_estimate) fits the HBM budget and a slot is free, gives it its own
solver on its own HIP stream, and keeps its state resident until it has reached
its target time.  Every active request advances in chunks of device-driven SSP-
RK3 steps on its own stream, so the work of all admitted requests overlaps on
the GPU; completion is detected without blocking (azb_loop_probe*).  Admission
is FIFO: a request that does not fit blocks the ones behind it, so large
requests are not starved.

azeban itself has no serving mode; each request is one Simulation::simulate_until
(src/azeban/simulation.cpp:89-102) of an IncompressibleEuler + SSP_RK3 solver.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from ._capi import SVParams, check, load
from .solver import PRECISIONS


@dataclass
class Request:
    rid: int
    dim: int
    N: int
    u_hat: torch.Tensor  # (dim, N, ..., N/2+1) complex128, host or device
    t_end: float
    arrival: float = 0.0  # seconds after the server starts
    precision: str = "fp64"


@dataclass
class _Active:
    req: Request
    handle: C.c_void_p
    stream: C.c_void_p
    u: torch.Tensor
    nbytes: int
    admitted: float
    steps: int = 0
    probing: bool = False
    # pinned host copy of a host-side initial state: kept alive until the request
    # finishes, since the asynchronous upload on the request's stream reads it
    host_src: torch.Tensor | None = None


@dataclass
class Result:
    rid: int
    u_hat: torch.Tensor
    steps: int
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class Server:
    hbm_budget: int = 200 * 2**30
    max_active: int = 32
    chunk_steps: int = 16
    C: float = 0.2
    results: list = field(default_factory=list)

    def run(self, requests: list[Request], poll_sleep: float = 2e-5) -> dict:
        lib = load()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        active: list[_Active] = []
        used = 0
        peak_active = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        while pending or active:
            # ---- admission (FIFO, by footprint)
            while pending and pending[0].arrival <= now() and len(active) < self.max_active:
                r = pending[0]
                need = int(lib.azb_bytes_estimate(r.dim, r.N, PRECISIONS[r.precision], 1))
                need += r.u_hat.numel() * 16
                if used + need > self.hbm_budget and active:
                    break
                pending.popleft()
                st = C.c_void_p()
                check(lib.azb_stream_create(C.byref(st)))
                h = C.c_void_p()
                visc = SVParams(0.05 / r.N, 1.0, 2 * math.pi * r.N / 4)
                check(lib.azb_create(r.dim, r.N, 0, r.dim, PRECISIONS[r.precision], visc, st, C.byref(h)))
                u = torch.empty(r.u_hat.shape, dtype=torch.complex128, device="cuda")
                src = r.u_hat.detach().to(torch.complex128)
                host_src = None
                if src.is_cuda:
                    torch.cuda.synchronize()
                    u.copy_(src)
                    torch.cuda.synchronize()
                else:
                    host_src = src.contiguous().pin_memory()
                    check(lib.azb_memcpy_h2d_async(C.c_void_p(u.data_ptr()), C.c_void_p(host_src.data_ptr()),
                                                   u.numel() * 16, st))
                check(lib.azb_loop_begin(h, r.t_end, self.C, None))
                a = _Active(r, h, st, u, need, now(), host_src=host_src)
                check(lib.azb_enqueue_steps(h, C.c_void_p(u.data_ptr()), self.chunk_steps))
                check(lib.azb_loop_probe(h))
                a.probing = True
                active.append(a)
                used += need
                peak_active = max(peak_active, len(active))
            # ---- progress
            progressed = False
            for a in list(active):
                ready = C.c_int(0)
                tt = C.c_double(0)
                ns = C.c_int(0)
                check(lib.azb_loop_probe_result(a.handle, C.byref(ready), C.byref(tt), C.byref(ns)))
                if not ready.value:
                    continue
                progressed = True
                a.steps = ns.value
                if tt.value < a.req.t_end:
                    check(lib.azb_enqueue_steps(a.handle, C.c_void_p(a.u.data_ptr()), self.chunk_steps))
                    check(lib.azb_loop_probe(a.handle))
                    continue
                check(lib.azb_stream_synchronize(a.stream))
                self.results.append(Result(a.req.rid, a.u, a.steps, a.req.arrival, a.admitted, now()))
                check(lib.azb_destroy(a.handle))
                check(lib.azb_stream_destroy(a.stream))
                active.remove(a)
                used -= a.nbytes
            if not progressed:
                if not active and pending:
                    wait = pending[0].arrival - now()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                else:
                    time.sleep(poll_sleep)
        wall = now()
        lat = np.array([r.latency for r in self.results])
        return {
            "requests": len(self.results),
            "wall_s": wall,
            "throughput_per_s": len(self.results) / wall if wall > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "peak_active": peak_active,
            "grid_point_steps": float(sum(r_.steps * _pts(r_) for r_ in self._reqs(requests))),
        }

    def _reqs(self, requests):
        by_id = {r.rid: r for r in requests}
        for res in self.results:
            yield _Done(by_id[res.rid], res.steps)


@dataclass
class _Done:
    req: Request
    steps: int


def _pts(d: _Done) -> float:
    return float(d.req.N) ** d.req.dim


def poisson_requests(n: int, rate: float, seed: int = 1, sizes_2d=(128, 256, 512, 1024),
                     sizes_3d=(64, 128, 256), t_end: float = 0.5, hurst: float = 0.75,
                     precision: str = "fp64") -> list[Request]:
    """Poisson arrivals (exponential gaps of mean 1/rate), sizes uniform over the given
    2D and 3D grids, fractional-Brownian initial data (|c_k| ~ |k|^-(H + d/2))."""
    from . import init as I
    from .solver import initialize

    rng = np.random.Generator(np.random.MT19937(seed))
    sizes = [(2, n_) for n_ in sizes_2d] + [(3, n_) for n_ in sizes_3d]
    out = []
    t = 0.0
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        dim, N = sizes[int(rng.integers(len(sizes)))]
        u = initialize(I.fractional_brownian(dim, N, seed=seed * 1000 + i, hurst=hurst))
        out.append(Request(i, dim, N, u, t_end, t, precision))
    torch.cuda.synchronize()
    return out
