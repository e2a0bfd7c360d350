"""Batched uneven-size ensembles (hipGraph replay) give the same per-sample
results as solving each sample alone, bit for bit."""
import numpy as np
import pytest
import torch

from azeban_amd import init as I
from azeban_amd import initialize
from azeban_amd.ensemble import Ensemble, Group, solve_one

pytestmark = pytest.mark.gpu


def dvp_batch(N, seeds):
    return torch.stack([initialize(I.vortex_patch(N, seed=s)) for s in seeds]).contiguous()


@pytest.mark.parametrize("use_graph", [True, False])
def test_ensemble_matches_single_solves_2d(cuda, use_graph):
    print(f"[case] 2D ensemble graph={use_graph}", flush=True)
    t_end = 0.02
    groups = [Group(2, 64, dvp_batch(64, [1, 2, 3])), Group(2, 128, dvp_batch(128, [4, 5]))]
    init = [g.u.clone() for g in groups]
    ens = Ensemble(groups)
    res = ens.run(t_end, steps_per_launch=4, use_graph=use_graph)
    torch.cuda.synchronize()
    for gi, g in enumerate(groups):
        for si in range(g.S):
            ref, n = solve_one(2, g.N, init[gi][si], t_end)
            assert res["steps"][gi][si] == n
            assert not (res["times"][gi][si] < t_end)
            assert torch.equal(g.u[si], ref), (gi, si)
    ens.close()


def test_ensemble_matches_single_solves_3d(cuda):
    print("[case] 3D ensemble", flush=True)
    t_end = 0.01
    u0 = torch.stack([initialize(I.fractional_brownian(3, 64, seed=s)) for s in (1, 2)]).contiguous()
    g = Group(3, 64, u0.clone())
    ens = Ensemble([g])
    res = ens.run(t_end, steps_per_launch=3)
    torch.cuda.synchronize()
    for si in range(2):
        ref, n = solve_one(3, 64, u0[si], t_end)
        assert res["steps"][0][si] == n
        assert torch.equal(g.u[si], ref)
    ens.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_ensemble_matches_single_solves_3d_256(cuda, precision):
    """3D batches at N = 256 (two line pairs per row-kernel workgroup, samples
    stacked in the launch) give each sample's single-solve result bit for bit."""
    print(f"[case] 3D ensemble N=256 {precision}", flush=True)
    t_end = 5e-3  # at least two steps: the viscous clamp caps dt at 3.2e-3
    u0 = torch.stack([initialize(I.fractional_brownian(3, 256, seed=s)) for s in (4, 5)]).contiguous()
    g = Group(3, 256, u0.clone(), precision)
    ens = Ensemble([g])
    res = ens.run(t_end, steps_per_launch=2)
    torch.cuda.synchronize()
    for si in range(2):
        ref, n = solve_one(3, 256, u0[si], t_end, precision=precision)
        assert n >= 2 and res["steps"][0][si] == n
        assert torch.equal(g.u[si], ref), si
    ens.close()


def test_graph_replays_repeatedly_match_eager(cuda):
    """Replaying the captured graph several times, with eager work and host syncs
    between replays, matches enqueuing the same steps eagerly."""
    print("[case] graph replay vs eager", flush=True)
    u0 = dvp_batch(64, [7, 8])
    a = Group(2, 64, u0.clone())
    b = Group(2, 64, u0.clone())
    ea, eb = Ensemble([a]), Ensemble([b])
    for rep in range(3):
        ra = ea.run(0.01 * (rep + 1), steps_per_launch=2, use_graph=True)
        rb = eb.run(0.01 * (rep + 1), steps_per_launch=2, use_graph=False)
        _ = torch.ones(10, device="cuda").sum().item()  # unrelated eager work between replays
        assert ra["steps"] == rb["steps"]
        assert torch.equal(a.u, b.u)
    ea.close()
    eb.close()


@pytest.mark.parametrize("dim,N", [(2, 1024), (2, 128), (3, 64)])
def test_autotuned_shapes_give_identical_results(cuda, dim, N):
    """Every launch shape the autotuner may pick computes each column / line pair
    with the same arithmetic, so results are bit-identical to the defaults."""
    import math

    import azeban_amd as A

    print(f"[case] autotune dim={dim} N={N}", flush=True)
    u = initialize(I.fractional_brownian(dim, N, seed=3))
    g = A.Grid(dim, N)
    eq = A.IncompressibleEuler(g, A.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4))
    ref = torch.empty_like(u)
    eq.dudt(ref, u)
    for ts, pd in [([2] * 6, 2), ([4] * 6, 2), ([4, 1, 2, 1, 4, 1], 1)]:
        eq.set_tuning(ts, pd)
        out = torch.empty_like(u)
        eq.dudt(out, u)
        assert torch.equal(out, ref), (ts, pd)
    rec = eq.autotune(u, reps=2)
    assert all(v in (1, 2, 4) for v in rec["tsdiv"]) and rec["pdiv"] in (1, 2)
    assert all(0 <= p < 3 for p in rec["plans"]) and rec["plans"][2] == rec["plans"][5] == 0
    # the tuned launch shapes with the default radix plans: identical to the defaults
    eq.set_plans([0] * 6)
    out = torch.empty_like(u)
    eq.dudt(out, u)
    assert torch.equal(out, ref)


def _solve_seq(dim, N, u0, targets, precision="fp64"):
    """One sample alone through Simulation, one simulate_until per target."""
    import math

    import azeban_amd as A

    g = A.Grid(dim, N)
    eq = A.IncompressibleEuler(g, A.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4), precision)
    sim = A.Simulation(g, 0.2, A.SSP_RK3(g.shape_fourier(dim), eq), u0)
    for t in targets:
        sim.simulate_until(t)
    return sim.u(), sim.time()


@pytest.mark.parametrize("use_graph", [True, False])
def test_second_run_continues_from_reached_time(cuda, use_graph):
    """run(a) then run(b) continues each sample from a, like simulate_until(a) then
    simulate_until(b) (src/azeban/simulation.cpp:89-102), bit for bit."""
    print(f"[case] ensemble run continuation graph={use_graph}", flush=True)
    groups = [Group(2, 64, dvp_batch(64, [1, 2])), Group(2, 128, dvp_batch(128, [3]))]
    init = [g.u.clone() for g in groups]
    ens = Ensemble(groups)
    ens.run(0.01, steps_per_launch=3, use_graph=use_graph)
    res = ens.run(0.025, steps_per_launch=3, use_graph=use_graph)
    torch.cuda.synchronize()
    for gi, g in enumerate(groups):
        for si in range(g.S):
            assert res["times"][gi][si] == 0.025
            ref, t = _solve_seq(2, g.N, init[gi][si], [0.01, 0.025])
            assert t == 0.025
            assert torch.equal(g.u[si], ref), (gi, si)
    # reset() starts over from time 0: one more run equals a fresh solve to that time
    ens.reset()
    for g, u0 in zip(groups, init):
        g.u.copy_(u0)
    ens.run(0.01, steps_per_launch=3, use_graph=use_graph)
    torch.cuda.synchronize()
    for gi, g in enumerate(groups):
        for si in range(g.S):
            ref, _ = _solve_seq(2, g.N, init[gi][si], [0.01])
            assert torch.equal(g.u[si], ref), (gi, si)
    ens.close()


def test_graph_recaptured_when_chunk_or_state_changes(cuda):
    """A replay never uses a stale chunk size or a replaced state tensor: changing
    either captures the graph again."""
    print("[case] ensemble graph re-capture", flush=True)
    u0 = dvp_batch(64, [5, 6])
    g = Group(2, 64, u0.clone())
    ens = Ensemble([g])
    ens.run(0.01, steps_per_launch=2)
    k1 = ens._graph_key
    ens.run(0.02, steps_per_launch=5)
    assert ens._graph_key != k1 and ens._graph_key[0] == 5
    old = g.u
    kept = old.clone()
    g.u = old.clone()  # the state moves to a new tensor
    ens.run(0.03, steps_per_launch=5)
    torch.cuda.synchronize()
    assert torch.equal(old, kept), "the replaced tensor was written"
    for si in range(2):
        ref, _ = _solve_seq(2, 64, u0[si], [0.01, 0.02, 0.03])
        assert torch.equal(g.u[si], ref), si
    ens.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_ensemble_config_sizes_against_single_solves_and_oracle(cuda, precision):
    """The ensemble workload's grid sizes (discontinuous vortex patches at 128^2, 256^2
    and 512^2, here 3/2/1 samples) in one graph, FP64 and FP32-mixed: every sample is
    bit-identical to its single solve in the same precision, and one sample per size
    matches the FP64 oracle (the sample loop of src/azeban/run_from_config.cpp:81-102)
    to relative L2 1e-11 (FP64) or 1e-5 (FP32-mixed)."""
    import numpy as np

    from oracle import azeban_oracle as O

    print(f"[case] ensemble config sizes {precision}", flush=True)
    t_end = 0.05
    seeds = {128: [11, 12, 13], 256: [21, 22], 512: [31]}
    groups = [Group(2, N, dvp_batch(N, s), precision) for N, s in seeds.items()]
    init = [g.u.clone() for g in groups]
    ens = Ensemble(groups)
    res = ens.run(t_end, steps_per_launch=4)
    torch.cuda.synchronize()
    tol = 1e-11 if precision == "fp64" else 1e-5
    for gi, g in enumerate(groups):
        for si in range(g.S):
            ref, n = solve_one(2, g.N, init[gi][si], t_end, precision=precision)
            assert res["steps"][gi][si] == n
            assert torch.equal(g.u[si], ref), (g.N, si)
        osim = O.Simulation(O.IncompressibleEuler(O.Grid(2, g.N), O.baseline_visc(g.N)),
                            init[gi][0].cpu().numpy(), 0.2)
        osim.simulate_until(t_end)
        err = O.l2_rel(g.u[0].cpu().numpy(), osim.u)
        print(f"  N={g.N} steps={res['steps'][gi][0]} oracle steps={len(osim.dts)} rel L2 {err:.2e}", flush=True)
        assert err <= tol, (g.N, err)
        if precision == "fp64":
            assert res["steps"][gi][0] == len(osim.dts)
    ens.close()
