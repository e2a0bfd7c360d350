"""The continuous-batching server: each request's result is the same alone and in
the mix (bit for bit), admission respects the HBM budget, and latencies are
reported."""
import pytest
import torch

from azeban_amd import init as I
from azeban_amd import initialize
from azeban_amd.ensemble import solve_one
from azeban_amd.serving import Request, Server

pytestmark = pytest.mark.gpu


def reqs():
    out = []
    specs = [(2, 64, 0.0), (2, 128, 0.001), (3, 64, 0.002), (2, 64, 0.003), (2, 256, 0.004), (3, 64, 0.005)]
    for i, (dim, N, arr) in enumerate(specs):
        u = initialize(I.fractional_brownian(dim, N, seed=100 + i))
        out.append(Request(i, dim, N, u, 0.004 if dim == 3 else 0.01, arr))
    torch.cuda.synchronize()
    return out


def test_results_same_alone_and_in_mix(cuda):
    print("[case] serving mix", flush=True)
    rs = reqs()
    srv = Server(chunk_steps=3, max_active=8)
    stats = srv.run(rs)
    torch.cuda.synchronize()
    assert stats["requests"] == len(rs)
    assert stats["peak_active"] >= 2  # work of several requests overlapped
    by_id = {r.rid: r for r in rs}
    for res in srv.results:
        r = by_id[res.rid]
        ref, n = solve_one(r.dim, r.N, r.u_hat, r.t_end)
        assert res.steps == n
        assert torch.equal(res.u_hat, ref), r.rid
    assert stats["latency_p99_s"] >= stats["latency_p50_s"] > 0


def test_budget_serialises_admission(cuda):
    """With a budget that fits one request at a time, requests still all complete
    (FIFO) and never overlap."""
    print("[case] serving budget", flush=True)
    rs = reqs()[:3]
    srv = Server(hbm_budget=1, chunk_steps=4)
    stats = srv.run(rs)
    assert stats["requests"] == 3
    assert stats["peak_active"] == 1


def test_host_side_requests_upload_asynchronously(cuda):
    """Requests whose state is in host memory are uploaded on their own stream from a
    pinned copy the server keeps until they finish; results equal those of the same
    requests given on the device."""
    print("[case] serving host inputs", flush=True)
    rs = reqs()
    host_rs = [Request(r.rid, r.dim, r.N, r.u_hat.cpu(), r.t_end, r.arrival) for r in rs]
    a, b = Server(chunk_steps=3, max_active=8), Server(chunk_steps=3, max_active=8)
    a.run(rs)
    b.run(host_rs)
    torch.cuda.synchronize()
    ra = {r.rid: r for r in a.results}
    for res in b.results:
        assert res.steps == ra[res.rid].steps
        assert torch.equal(res.u_hat, ra[res.rid].u_hat), res.rid
