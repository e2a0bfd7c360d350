"""The C-ABI library loads and exports every symbol include/azeban_hip.h declares
(no compute call: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

from azeban_amd import _capi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "azeban_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(azb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for name in ("azb_create", "azb_dudt", "azb_rk3_step", "azb_simulate_until", "azb_fft_r2c",
                 "azb_fft_c2r", "azb_copy_to_padded", "azb_axpby", "azb_max_norm"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_capi.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_capi.SIGNATURES)


def test_load_sets_prototypes():
    lib = _capi.load()
    assert lib.azb_last_error() is not None


# ---------------------------------------------------------------------------
# Argument validation: these entry points reject bad input with a status code and
# a message before any device call (the reference's LOG_ERR checks become codes)
# ---------------------------------------------------------------------------
import ctypes as C  # noqa: E402

import pytest  # noqa: E402


def _create(dim, N, M=0, prec=None, samples=1):
    lib = _capi.load()
    h = C.c_void_p()
    visc = _capi.SVParams(0.05 / max(N, 1), 1.0, 1.0)
    prec = _capi.AZB_FP64 if prec is None else prec
    rc = lib.azb_create_batch(dim, N, M, dim, prec, visc, samples, None, C.byref(h))
    return rc, h, lib.azb_last_error().decode()


@pytest.mark.parametrize("dim,N,M,prec,samples,code,needle", [
    (4, 64, 0, None, 1, -1, "dim"),
    (3, 66, 0, None, 1, -1, "multiple of 4"),
    (3, 64, 100, None, 1, -3, "3N/2"),
    (3, 100, 0, None, 1, -3, "not built"),
    (3, 1024, 0, None, 1, -3, "not built"),
    (2, 64, 0, 7, 1, -1, "precision"),
    (2, 64, 0, None, 0, -1, "nsamples"),
])
def test_create_rejects_bad_arguments(dim, N, M, prec, samples, code, needle):
    rc, h, msg = _create(dim, N, M, prec, samples)
    assert rc == code and not h.value
    assert needle in msg


def test_entry_points_reject_null_handles():
    lib = _capi.load()
    dt = C.c_double()
    assert lib.azb_dudt(None, None, None) == _capi.AZB_ERR_ARG
    assert lib.azb_rk3_step(None, None, 1.0, 0.2, C.byref(dt)) == _capi.AZB_ERR_ARG
    assert lib.azb_dt(None, 0.2, C.byref(dt)) == _capi.AZB_ERR_ARG
    assert lib.azb_set_tuning(None, None, 1) == _capi.AZB_ERR_ARG
    assert lib.azb_copy_bandwidth(-1.0, 1, C.byref(dt)) == _capi.AZB_ERR_ARG
    assert lib.azb_copy_from_padded(2, 24, 16, None, None, None) == _capi.AZB_ERR_ARG
    assert lib.azb_copy_from_padded(4, 24, 16, C.c_void_p(8), C.c_void_p(8), None) == _capi.AZB_ERR_ARG
    assert lib.azb_scale(1.0, None, 4, None) == _capi.AZB_ERR_ARG
    assert "NULL" in lib.azb_last_error().decode() or lib.azb_last_error()


def test_transform_and_helper_entry_points_reject_bad_arguments():
    """Argument checks that return before any device call (no GPU needed)."""
    lib = _capi.load()
    p = C.c_void_p(8)
    out = C.c_double()
    spec = (C.c_double * 8)()
    for fn in (lib.azb_fft_r2c, lib.azb_fft_c2r):
        assert fn(3, 64, 1, 0, None, p, None) == _capi.AZB_ERR_ARG
        assert fn(3, 64, 1, 7, p, p, None) == _capi.AZB_ERR_ARG
        assert "precision" in lib.azb_last_error().decode()
        assert fn(3, 64, 0, 0, p, p, None) == _capi.AZB_ERR_ARG
        assert fn(3, 100, 1, 0, p, p, None) == _capi.AZB_ERR_UNSUPPORTED
        assert fn(4, 64, 1, 0, p, p, None) == _capi.AZB_ERR_ARG
    assert lib.azb_leray(2, 64, None, None) == _capi.AZB_ERR_ARG
    assert lib.azb_leray(4, 64, p, None) == _capi.AZB_ERR_ARG
    assert lib.azb_leray(2, 0, p, None) == _capi.AZB_ERR_ARG
    assert lib.azb_energy_spectrum(2, 64, None, spec, None) == _capi.AZB_ERR_ARG
    assert lib.azb_energy_spectrum(4, 64, p, spec, None) == _capi.AZB_ERR_ARG
    assert lib.azb_energy_spectrum(2, 7, p, spec, None) == _capi.AZB_ERR_ARG
    assert lib.azb_max_norm(None, 4, C.byref(out), None) == _capi.AZB_ERR_ARG
    assert lib.azb_axpby(1.0, None, 1.0, p, 4, None) == _capi.AZB_ERR_ARG
