"""Continuous-batching solve server.

Requests (a 2D or 3D state of any built size plus a target time) arrive over
time.  For every (dim, N, precision) class the server keeps one batched solver
(azb_create_batch) with a fixed number of request slots on its own HIP stream.
A waiting request is admitted into a free slot of its class as soon as one is
free and, for a class that has no solver yet, as soon as that solver's device
footprint (azb_bytes_estimate) fits the HBM budget (idle classes are released to
make room).  Admission queues the upload of the request's state into its slot
and a per-slot restart of the device-driven loop (azb_loop_begin_sample) on the
class stream, behind the chunks already queued there: requests join and leave a
running batch without stopping it.  Every launch of a class covers all of its
slots (free and finished slots return at once), the classes run on separate
streams, and nothing in the loop blocks the host or the device: progress comes
from non-blocking probes (azb_loop_probe*), and two chunks stay queued per class
so the device runs one while the host reads the progress of the other.
Admission is FIFO: a request that does not fit blocks the ones behind it, so
large requests are not starved.

azeban itself has no serving mode; each request is one Simulation::simulate_until
(src/azeban/simulation.cpp:89-102) of an IncompressibleEuler + SSP_RK3 solver,
and its result is bit-identical to solving it alone.
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


# default request slots per class: enough concurrent requests of one size to fill
# the device with one launch per kernel, few enough to keep idle slots cheap
DEFAULT_SLOTS = {(2, 64): 16, (2, 128): 16, (2, 256): 8, (2, 512): 4, (2, 1024): 2,
                 (3, 64): 8, (3, 128): 2, (3, 256): 1, (3, 512): 1}


@dataclass
class _Slot:
    req: Request
    admitted: float
    first_probe: int  # the first probe queued after this request's restart of the slot
    host_src: torch.Tensor | None = None  # upload source, alive until the request ends


class _Class:
    """One batched solver of a (dim, N, precision) class and its request slots."""

    def __init__(self, lib, dim: int, N: int, precision: str, slots: int, C_: float):
        self.lib = lib
        self.key = (dim, N, precision)
        self.stream = C.c_void_p()
        check(lib.azb_stream_create(C.byref(self.stream)))
        self.ext = torch.cuda.ExternalStream(self.stream.value)
        self.handle = C.c_void_p()
        visc = SVParams(0.05 / N, 1.0, 2 * math.pi * N / 4)
        rc = lib.azb_create_batch(dim, N, 0, dim, PRECISIONS[precision], visc, slots, self.stream,
                                  C.byref(self.handle))
        if rc:
            lib.azb_stream_destroy(self.stream)
            check(rc)
        shape = (slots, dim) + (N,) * (dim - 1) + (N // 2 + 1,)
        self.u = torch.zeros(shape, dtype=torch.complex128, device="cuda")
        # the slots start finished (time 0 = target 0): launches skip them
        check(lib.azb_loop_begin(self.handle, 0.0, C_, None))
        self.slots: list[_Slot | None] = [None] * slots
        self.nbytes = int(lib.azb_bytes_estimate(dim, N, PRECISIONS[precision], slots)) + self.u.numel() * 16
        self.probing = False  # a progress probe is queued and not read yet
        self.probe_seq = 0  # number of probes queued so far
        self.last_used = 0.0
        self._t = (C.c_double * slots)()
        self._n = (C.c_int * slots)()

    def active(self) -> int:
        return sum(s is not None for s in self.slots)

    def free_slot(self) -> int:
        for i, s in enumerate(self.slots):
            if s is None:
                return i
        return -1

    def chunk(self, steps: int):
        check(self.lib.azb_enqueue_steps(self.handle, C.c_void_p(self.u.data_ptr()), steps))

    def probe(self):
        check(self.lib.azb_loop_probe(self.handle))
        self.probing = True
        self.probe_seq += 1

    def probe_result(self):
        ready = C.c_int(0)
        check(self.lib.azb_loop_probe_result(self.handle, C.byref(ready), self._t, self._n))
        if not ready.value:
            return None
        self.probing = False
        return list(self._t), list(self._n)

    def close(self):
        if self.handle:
            self.lib.azb_destroy(self.handle)  # waits for the class stream
            self.handle = None
        self.u = None
        self.ext = None
        if self.stream:
            self.lib.azb_stream_destroy(self.stream)
            self.stream = None


@dataclass
class Server:
    hbm_budget: int = 200 * 2**30
    max_active: int = 64
    chunk_steps: int = 16
    C: float = 0.2
    slots: dict = field(default_factory=dict)  # (dim, N) -> slots, overriding DEFAULT_SLOTS
    stall_s: float = 600.0  # running classes without any progress for this long is an error
    results: list = field(default_factory=list)

    def _slots_for(self, dim: int, N: int) -> int:
        return int(self.slots.get((dim, N), DEFAULT_SLOTS.get((dim, N), 1)))

    def run(self, requests: list[Request], poll_sleep: float = 2e-5) -> dict:
        lib = load()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        classes: dict[tuple, _Class] = {}
        peak_active = 0
        peak_bytes = 0
        launches = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def used():
            return sum(c.nbytes for c in classes.values())

        def n_active():
            return sum(c.active() for c in classes.values())

        def make_room(need: int, keep) -> bool:
            """Release idle classes (least recently used first) until `need` more bytes fit."""
            idle = sorted((c for c in classes.values() if c.active() == 0 and c.key != keep),
                          key=lambda c: c.last_used)
            for c in idle:
                if used() + need <= self.hbm_budget:
                    break
                c.close()
                del classes[c.key]
            return used() + need <= self.hbm_budget

        last_progress = 0.0
        try:
            while pending or n_active():
                # ---- admission (FIFO): a free slot of the request's class, creating the
                # class when its footprint fits the budget
                while pending and pending[0].arrival <= now() and n_active() < self.max_active:
                    r = pending[0]
                    key = (r.dim, r.N, r.precision)
                    cl = classes.get(key)
                    if cl is None:
                        S = self._slots_for(r.dim, r.N)
                        need = (int(lib.azb_bytes_estimate(r.dim, r.N, PRECISIONS[r.precision], S))
                                + S * r.u_hat.numel() * 16)
                        if not make_room(need, key) and n_active():
                            break
                        cl = classes[key] = _Class(lib, r.dim, r.N, r.precision, S, self.C)
                        peak_bytes = max(peak_bytes, used())
                    i = cl.free_slot()
                    if i < 0:
                        break
                    pending.popleft()
                    self._admit(lib, cl, i, r, now())
                    last_progress = now()
                    if not cl.probing:
                        # the class was idle: start it with two chunks, probed after the first
                        cl.chunk(self.chunk_steps)
                        cl.probe()
                        cl.chunk(self.chunk_steps)
                        launches += 2
                    peak_active = max(peak_active, n_active())
                # ---- progress of every running class
                progressed = False
                for cl in list(classes.values()):
                    if not cl.probing:
                        continue
                    st = cl.probe_result()
                    if st is None:
                        continue
                    progressed = True
                    times, steps = st
                    for i, s in enumerate(cl.slots):
                        # a probe queued before the slot's restart still shows the slot's
                        # previous request
                        if s is None or s.first_probe > cl.probe_seq or times[i] < s.req.t_end:
                            continue
                        self._finish(cl, i, steps[i], now())
                    cl.last_used = now()
                    if cl.active():
                        cl.probe()  # behind the chunk already queued
                        cl.chunk(self.chunk_steps)
                        launches += 1
                if progressed:
                    last_progress = now()
                elif n_active() and now() - last_progress > self.stall_s:
                    raise RuntimeError(f"no request made progress for {self.stall_s} s")
                if not progressed:
                    if not n_active() and pending:
                        wait = pending[0].arrival - now()
                        if wait > 0:
                            time.sleep(min(wait, 0.01))
                    else:
                        time.sleep(poll_sleep)
            wall = now()
        finally:
            for cl in classes.values():
                cl.close()
        torch.cuda.synchronize()
        return self._stats(requests, wall, peak_active, peak_bytes, launches)

    def _admit(self, lib, cl: _Class, i: int, r: Request, t: float):
        """Queue the request's state into slot i and restart that slot's loop, on the class
        stream behind the work already queued there (neither host nor device waits)."""
        src = r.u_hat.detach()
        want = (r.dim,) + (r.N,) * (r.dim - 1) + (r.N // 2 + 1,)
        if tuple(src.shape) != want:
            raise ValueError(f"request {r.rid}: state shape {tuple(src.shape)}, expected {want}")
        host_src = None
        if src.is_cuda:
            src = src.to(torch.complex128).contiguous()
            cl.ext.wait_stream(torch.cuda.current_stream())  # the state was made on torch's stream
            with torch.cuda.stream(cl.ext):
                cl.u[i].copy_(src)
            host_src = src  # kept until the request ends, like a host source
        else:
            host_src = src.to(torch.complex128).contiguous().pin_memory()
            check(lib.azb_memcpy_h2d_async(C.c_void_p(cl.u[i].data_ptr()), C.c_void_p(host_src.data_ptr()),
                                           cl.u[i].numel() * 16, cl.stream))
        check(lib.azb_loop_begin_sample(cl.handle, i, r.t_end, self.C, 0.0))
        cl.slots[i] = _Slot(r, t, cl.probe_seq + 1, host_src)

    def _finish(self, cl: _Class, i: int, steps: int, t: float):
        """Copy slot i's state out on the class stream (ordered before any later upload
        into the slot) and hand it to torch's stream."""
        s = cl.slots[i]
        out = torch.empty_like(cl.u[i])
        with torch.cuda.stream(cl.ext):
            out.copy_(cl.u[i])
        torch.cuda.current_stream().wait_stream(cl.ext)
        self.results.append(Result(s.req.rid, out, steps, s.req.arrival, s.admitted, t))
        cl.slots[i] = None

    def _stats(self, requests, wall, peak_active, peak_bytes, launches) -> dict:
        by_id = {r.rid: r for r in requests}
        lat = np.array([r.latency for r in self.results])
        return {
            "requests": len(self.results),
            "wall_s": wall,
            "throughput_per_s": len(self.results) / wall if wall > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "peak_active": peak_active,
            "peak_solver_bytes": peak_bytes,
            "chunk_launches": launches,
            "grid_point_steps": float(sum(res.steps * float(by_id[res.rid].N) ** by_id[res.rid].dim
                                          for res in self.results)),
        }


def serve_one_at_a_time(requests: list[Request], C_: float = 0.2, chunk: int = 16) -> dict:
    """The baseline the server is compared with: the same requests in arrival order,
    each solved alone (one IncompressibleEuler per grid size, created before the clock
    starts and reused) once it has arrived and the previous one has finished."""
    from .ensemble import solve_one
    from .solver import Grid, IncompressibleEuler, SmoothCutoff1D

    eqs = {}
    for r in requests:
        key = (r.dim, r.N, r.precision)
        if key not in eqs:
            eqs[key] = IncompressibleEuler(Grid(r.dim, r.N), SmoothCutoff1D(0.05 / r.N, 1.0, 2 * math.pi * r.N / 4),
                                           r.precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = []
    pts = 0.0
    for r in sorted(requests, key=lambda r: r.arrival):
        wait = r.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        start = time.perf_counter() - t0
        u, n = solve_one(r.dim, r.N, r.u_hat, r.t_end, C_, r.precision, chunk, eq=eqs[(r.dim, r.N, r.precision)])
        torch.cuda.synchronize()
        results.append(Result(r.rid, u, n, r.arrival, start, time.perf_counter() - t0))
        pts += n * float(r.N) ** r.dim
    wall = time.perf_counter() - t0
    lat = np.array([x.latency for x in results])
    return {"requests": len(results), "wall_s": wall, "throughput_per_s": len(results) / wall,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "grid_point_steps": pts, "results": results}


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
