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
