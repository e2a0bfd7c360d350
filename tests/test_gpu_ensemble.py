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
    out = torch.empty_like(u)
    eq.dudt(out, u)
    assert torch.equal(out, ref)
