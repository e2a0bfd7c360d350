"""Batched ensembles of uneven grid sizes on one GPU.

azeban runs the samples of a statistical-solution ensemble one after another
(src/azeban/run_from_config.cpp:81-102), one MPI sample group per rank
(src/azeban.cpp:136-152).  Here the samples of each grid size form one batched
solver (azb_create_batch: every kernel launch covers all of them, each sample
keeps its own dt/time on the device), all size groups share one stream, and a
chunk of steps of every group is captured once into a hipGraph and replayed
until every sample has reached the target time.  Samples that finish early turn
their share of each launch into an immediate return.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import torch

from ._capi import SVParams, check, load
from .solver import PRECISIONS


def _vp(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


@dataclass
class Group:
    """All samples of one grid size: state (S, dim, N, ..., N/2+1) complex128."""

    dim: int
    N: int
    u: torch.Tensor
    precision: str = "fp64"
    handle: C.c_void_p = field(default=None, repr=False)

    @property
    def S(self) -> int:
        return self.u.shape[0]


class Ensemble:
    def __init__(self, groups: list[Group], C_: float = 0.2, eps_factor: float = 0.05):
        self.lib = load()
        self.C = C_
        st = C.c_void_p()
        check(self.lib.azb_stream_create(C.byref(st)))
        self.stream = st
        self.groups = groups
        for g in groups:
            h = C.c_void_p()
            visc = SVParams(eps_factor / g.N, 1.0, 2 * math.pi * g.N / 4)
            check(self.lib.azb_create_batch(g.dim, g.N, 0, g.dim, PRECISIONS[g.precision], visc, g.S,
                                            self.stream, C.byref(h)))
            g.handle = h
        self.graph = None

    def autotune(self, reps: int = 3, cache: str | None = None) -> list:
        """Pick each group's launch shapes (column tile widths, row line pairs) by timing
        them on that group's real batch; `cache` stores/reuses the choice per
        (dim, N, precision, samples)."""
        from .solver import _tuning_load, _tuning_save
        from ._capi import PROF_FAMILIES

        store = _tuning_load(cache) if cache else {}
        out = []
        for g in self.groups:
            key = f"{g.dim}/{g.N}/{g.precision}/{g.S}"
            if key in store:
                rec = store[key]
            else:
                torch.cuda.synchronize()
                check(self.lib.azb_autotune(g.handle, _vp(g.u), reps))
                ts = (C.c_int * len(PROF_FAMILIES))()
                pd = C.c_int()
                check(self.lib.azb_get_tuning(g.handle, ts, C.byref(pd)))
                rec = {"tsdiv": list(ts), "pdiv": pd.value}
                store[key] = rec
            ts = (C.c_int * len(PROF_FAMILIES))(*rec["tsdiv"])
            check(self.lib.azb_set_tuning(g.handle, ts, rec["pdiv"]))
            out.append(rec)
        if cache:
            _tuning_save(cache, store)
        if self.graph is not None:  # launch shapes changed: capture again
            self.lib.azb_graph_destroy(self.graph)
            self.graph = None
        return out

    def close(self):
        if self.graph is not None:
            self.lib.azb_graph_destroy(self.graph)
            self.graph = None
        for g in self.groups:
            if g.handle:
                self.lib.azb_destroy(g.handle)
                g.handle = None
        if self.stream:
            self.lib.azb_stream_destroy(self.stream)
            self.stream = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _times(self):
        out = []
        for g in self.groups:
            t = (C.c_double * g.S)()
            n = (C.c_int * g.S)()
            check(self.lib.azb_loop_read(g.handle, t, n))
            out.append((list(t), list(n)))
        return out

    def run(self, t_end: float, steps_per_launch: int = 16, use_graph: bool = True,
            max_launches: int = 100000) -> dict:
        """Integrate every sample to t_end; returns step counts and wall time."""
        torch.cuda.synchronize()
        for g in self.groups:
            check(self.lib.azb_loop_begin(g.handle, t_end, self.C, None))
        t0 = time.perf_counter()
        if use_graph and self.graph is None:
            check(self.lib.azb_capture_begin(self.stream))
            for g in self.groups:
                check(self.lib.azb_enqueue_steps(g.handle, _vp(g.u), steps_per_launch))
            gr = C.c_void_p()
            check(self.lib.azb_capture_end(self.stream, C.byref(gr)))
            self.graph = gr
        launches = 0
        while launches < max_launches:
            if use_graph:
                check(self.lib.azb_graph_launch(self.graph, self.stream))
            else:
                for g in self.groups:
                    check(self.lib.azb_enqueue_steps(g.handle, _vp(g.u), steps_per_launch))
            launches += 1
            status = self._times()
            if all(not (t < t_end) for times, _ in status for t in times):
                break
        wall = time.perf_counter() - t0
        status = self._times()
        return {"wall_s": wall, "launches": launches,
                "steps": [steps for _, steps in status], "times": [times for times, _ in status]}


def solve_one(dim: int, N: int, u_hat: torch.Tensor, t_end: float, C_: float = 0.2,
              precision: str = "fp64", chunk: int = 16, eq=None) -> tuple[torch.Tensor, int]:
    """One sample alone through the single-solve path (azb_create + azb_simulate_until);
    `eq` reuses an existing IncompressibleEuler of that grid (its workspace and tuning)."""
    from .solver import IncompressibleEuler, Grid, SmoothCutoff1D, Simulation, SSP_RK3

    g = Grid(dim, N)
    if eq is None:
        eq = IncompressibleEuler(g, SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4), precision)
    sim = Simulation(g, C_, SSP_RK3(g.shape_fourier(dim), eq), u_hat)
    n = sim.simulate_until(t_end, chunk=chunk)
    return sim.u(), n
