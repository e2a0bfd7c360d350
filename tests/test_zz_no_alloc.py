"""azb_dudt_noalloc: the RHS without a solver handle and without allocation (its
intermediates live in a file-scope device array).  On the GPU it must give exactly
azb_dudt's result on consecutive calls, the later ones on smaller grids (the array
keeps the previous call's contents), in place too; on the CPU, grids that do not fit
the array are rejected before any device call."""
import ctypes as C
import math

import pytest
import torch

from azeban_amd import _capi

PREC = {"fp64": _capi.AZB_FP64, "fp32_mixed": _capi.AZB_FP32_MIXED}


def _visc(N):
    return _capi.SVParams(0.05 / N, 1.0, 2 * math.pi * N / 4)


def test_rejects_what_does_not_fit_before_any_device_call():
    lib = _capi.load()
    one = C.c_void_p(1)  # never dereferenced: every case fails its argument checks first
    rc = lib.azb_dudt_noalloc(3, 128, _capi.AZB_FP64, _visc(128), one, one, None)
    assert rc == _capi.AZB_ERR_UNSUPPORTED and "scratch" in lib.azb_last_error().decode()
    rc = lib.azb_dudt_noalloc(3, 256, _capi.AZB_FP32_MIXED, _visc(256), one, one, None)
    assert rc == _capi.AZB_ERR_UNSUPPORTED
    rc = lib.azb_dudt_noalloc(2, 2048, _capi.AZB_FP64, _visc(2048), one, one, None)
    assert rc == _capi.AZB_ERR_UNSUPPORTED
    rc = lib.azb_dudt_noalloc(2, 64, 7, _visc(64), one, one, None)
    assert rc == _capi.AZB_ERR_ARG
    rc = lib.azb_dudt_noalloc(2, 64, _capi.AZB_FP64, _visc(64), None, one, None)
    assert rc == _capi.AZB_ERR_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_noalloc_equals_dudt_on_consecutive_calls(cuda, precision):
    import azeban_amd as A
    from azeban_amd import init as I
    from azeban_amd import initialize

    lib = _capi.load()
    # each call after the first runs on a smaller grid than the one before it
    for i, (dim, N) in enumerate([(2, 1024), (2, 128), (3, 64), (2, 64)]):
        print(f"[case] noalloc {precision} call {i}: {dim}D N={N}", flush=True)
        u = initialize(I.fractional_brownian(dim, N, seed=40 + i))
        eq = A.IncompressibleEuler(A.Grid(dim, N), A.SmoothCutoff1D(0.05 / N, 1.0, 2 * math.pi * N / 4),
                                   precision)
        ref = torch.empty_like(u)
        eq.dudt(ref, u)
        out = torch.full_like(u, float("nan"))
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        _capi.check(lib.azb_dudt_noalloc(dim, N, PREC[precision], _visc(N), C.c_void_p(u.data_ptr()),
                                         C.c_void_p(out.data_ptr()), st))
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (dim, N)
        # in place, like IncompressibleEuler::dudt with dudt == u
        w = u.clone()
        _capi.check(lib.azb_dudt_noalloc(dim, N, PREC[precision], _visc(N), C.c_void_p(w.data_ptr()),
                                         C.c_void_p(w.data_ptr()), st))
        torch.cuda.synchronize()
        assert torch.equal(w, ref), (dim, N, "in place")
