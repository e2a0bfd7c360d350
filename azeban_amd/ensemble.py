"""Batched ensembles of uneven grid sizes on one GPU.

azeban runs the samples of a statistical-solution ensemble one after another
(src/azeban/run_from_config.cpp:81-102), one MPI sample group per rank
(src/azeban.cpp:136-152).  Here the samples of each grid size form one batched
solver (azb_create_batch: every kernel launch covers all of them, each sample
keeps its own dt/time on the device) on its own stream, and a chunk of steps of
every group is captured once into a hipGraph (one branch per group) and replayed
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
    """Every size group gets its own solver on its own HIP stream; a chunk of steps of all
    groups is captured into one hipGraph whose branches (one per group) may run side by
    side on the device.

    Like Simulation::simulate_until (src/azeban/simulation.cpp:89-102), run(t) continues
    every sample from the time the previous run() left it at (t for a sample that got
    there), so run(a) followed by run(b) equals simulate_until(a) then simulate_until(b)
    of each sample alone.  reset() / set_times() mirror Simulation::reset / set_time.
    """

    def __init__(self, groups: list[Group], C_: float = 0.2, eps_factor: float = 0.05):
        self.lib = load()
        self.C = C_
        self.stream = self._new_stream()
        self.groups = groups
        self.streams = []
        for g in groups:
            gs = self._new_stream()
            self.streams.append(gs)
            h = C.c_void_p()
            visc = SVParams(eps_factor / g.N, 1.0, 2 * math.pi * g.N / 4)
            check(self.lib.azb_create_batch(g.dim, g.N, 0, g.dim, PRECISIONS[g.precision], visc, g.S,
                                            gs, C.byref(h)))
            g.handle = h
        self.graph = None
        self._graph_key = None
        self._t0 = [[0.0] * g.S for g in groups]  # start time of every sample in the next run()
        self.stall_s = 600.0  # a chunk that has not finished after this long is an error

    def _new_stream(self) -> C.c_void_p:
        st = C.c_void_p()
        check(self.lib.azb_stream_create(C.byref(st)))
        return st

    def times(self) -> list[list[float]]:
        """The time every sample has reached (what the next run() starts from)."""
        return [list(t) for t in self._t0]

    def set_times(self, times: list[list[float]]) -> None:
        if [len(t) for t in times] != [g.S for g in self.groups]:
            raise ValueError("set_times needs one time per sample of every group")
        self._t0 = [[float(x) for x in t] for t in times]

    def reset(self) -> None:
        self._t0 = [[0.0] * g.S for g in self.groups]

    def autotune(self, reps: int = 3, cache: str | None = None) -> list:
        """Pick each group's radix plans (once per grid size and precision, shared with
        single solves of that grid, so per-sample results stay identical to them) and
        launch shapes (column tile widths, row line pairs, per samples count) by timing
        them on that group's real batch; `cache` stores/reuses the choices."""
        from .solver import _ensure_plans, _tuning_load, _tuning_save
        from ._capi import PROF_FAMILIES

        out = []
        for g in self.groups:
            torch.cuda.synchronize()
            # radix plans per grid size, shared with single solves of that grid
            plans = _ensure_plans(g.handle, g.dim, g.N, g.precision, g.u, reps, cache)
            store = _tuning_load(cache) if cache else {}
            key = f"{g.dim}/{g.N}/{g.precision}/{g.S}"
            if key in store:
                rec = store[key]
            else:
                check(self.lib.azb_autotune(g.handle, _vp(g.u), reps))
                ts = (C.c_int * len(PROF_FAMILIES))()
                pd = C.c_int()
                check(self.lib.azb_get_tuning(g.handle, ts, C.byref(pd)))
                rec = {"tsdiv": list(ts), "pdiv": pd.value}
                if cache:
                    store[key] = rec
                    _tuning_save(cache, store)
            ts = (C.c_int * len(PROF_FAMILIES))(*rec["tsdiv"])
            check(self.lib.azb_set_tuning(g.handle, ts, rec["pdiv"]))
            out.append(dict(rec, plans=plans))
        self._drop_graph()  # launch shapes changed: capture again
        return out

    def _drop_graph(self):
        if self.graph is not None:
            self.lib.azb_graph_destroy(self.graph)
            self.graph = None
            self._graph_key = None

    def close(self):
        self._drop_graph()
        for g in self.groups:
            if g.handle:
                self.lib.azb_destroy(g.handle)
                g.handle = None
        for st in self.streams + [self.stream]:
            if st:
                self.lib.azb_stream_destroy(st)
        self.streams = []
        self.stream = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _enqueue_chunk(self, steps: int):
        """One chunk on every group's stream, forked from and joined back to the main
        stream (inside a capture this makes one graph branch per group)."""
        for g, gs in zip(self.groups, self.streams):
            check(self.lib.azb_stream_wait(gs, self.stream))
            check(self.lib.azb_enqueue_steps(g.handle, _vp(g.u), steps))
            check(self.lib.azb_stream_wait(self.stream, gs))

    def _launch(self, steps: int, use_graph: bool):
        if use_graph:
            check(self.lib.azb_graph_launch(self.graph, self.stream))
        else:
            self._enqueue_chunk(steps)

    def _probe(self):
        """Queue a copy of every sample's progress behind the chunks queued so far."""
        for g, gs in zip(self.groups, self.streams):
            check(self.lib.azb_stream_wait(gs, self.stream))
            check(self.lib.azb_loop_probe(g.handle))

    def _probe_wait(self, poll_sleep: float):
        """Per-group (times, steps) of the latest probes; the host polls, it never blocks a stream."""
        out = []
        for g in self.groups:
            t = (C.c_double * g.S)()
            n = (C.c_int * g.S)()
            ready = C.c_int(0)
            t0 = time.perf_counter()
            while True:
                check(self.lib.azb_loop_probe_result(g.handle, C.byref(ready), t, n))
                if ready.value:
                    break
                if time.perf_counter() - t0 > self.stall_s:
                    raise RuntimeError(f"no progress of group {g.dim}D N={g.N} for {self.stall_s} s")
                time.sleep(poll_sleep)
            out.append((list(t), list(n)))
        return out

    def run(self, t_end: float, steps_per_launch: int = 16, use_graph: bool = True,
            max_launches: int = 100000, poll_sleep: float = 5e-5) -> dict:
        """Integrate every sample from its current time to t_end; returns per-sample step
        counts of this call, the times reached and the wall time.  Two chunks stay queued
        at once: while the host reads the progress of one, the device runs the next (a
        chunk past t_end does nothing on the device)."""
        if steps_per_launch < 1:
            raise ValueError("steps_per_launch must be >= 1")
        torch.cuda.synchronize()
        for g, t0 in zip(self.groups, self._t0):
            check(self.lib.azb_loop_begin(g.handle, t_end, self.C, (C.c_double * g.S)(*t0)))
        t0_wall = time.perf_counter()
        key = (steps_per_launch, tuple(g.u.data_ptr() for g in self.groups))
        if use_graph and self._graph_key != key:
            self._drop_graph()
            check(self.lib.azb_capture_begin(self.stream))
            try:
                self._enqueue_chunk(steps_per_launch)
            finally:
                gr = C.c_void_p()
                rc = self.lib.azb_capture_end(self.stream, C.byref(gr))
            check(rc)
            self.graph, self._graph_key = gr, key
        self._launch(steps_per_launch, use_graph)
        self._probe()
        launches = 1
        if max_launches > 1:
            self._launch(steps_per_launch, use_graph)
            launches = 2
        while True:
            # progress as of the second-latest chunk; the latest one keeps the device busy
            status = self._probe_wait(poll_sleep)
            if all(not (t < t_end) for times, _ in status for t in times) or launches >= max_launches:
                break
            self._probe()  # behind the latest chunk, read on the next iteration
            self._launch(steps_per_launch, use_graph)
            launches += 1
        for gs in self.streams + [self.stream]:
            check(self.lib.azb_stream_synchronize(gs))
        wall = time.perf_counter() - t0_wall
        status = []
        for g in self.groups:
            t = (C.c_double * g.S)()
            n = (C.c_int * g.S)()
            check(self.lib.azb_loop_read(g.handle, t, n))
            status.append((list(t), list(n)))
        # a sample that reached t_end is at t_end exactly, as Simulation sets time_ = t
        self._t0 = [[t_end if not (t < t_end) else t for t in times] for times, _ in status]
        return {"wall_s": wall, "launches": launches,
                "steps": [steps for _, steps in status], "times": self.times()}


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
