"""ctypes binding of the C ABI in include/azeban_hip.h (libazeban_hip.so).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libazeban_hip.so"

AZB_FP64 = 0
AZB_FP32_MIXED = 1
AZB_OK = 0
AZB_ERR_ARG = -1
AZB_ERR_HIP = -2
AZB_ERR_UNSUPPORTED = -3
AZB_ERR_ALLOC = -4

PROF_FAMILIES = ["inv_ax0", "inv_ax1", "row_fused", "fwd_ax1", "fwd_ax0", "kspace_rk"]


class AzbError(RuntimeError):
    pass


class SVParams(C.Structure):
    _fields_ = [("eps", C.c_double), ("s", C.c_double), ("k0", C.c_double)]


_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_l = C.c_long
_sz = C.c_size_t
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# name -> (restype, argtypes); the full list of exported entry points
SIGNATURES = {
    "azb_last_error": (C.c_char_p, []),
    "azb_create": (_i, [_i, _i, _i, _i, _i, SVParams, _vp, C.POINTER(_vp)]),
    "azb_create_batch": (_i, [_i, _i, _i, _i, _i, SVParams, _i, _vp, C.POINTER(_vp)]),
    "azb_destroy": (_i, [_vp]),
    "azb_num_samples": (_i, [_vp]),
    "azb_bytes_estimate": (_sz, [_i, _i, _i, _i]),
    "azb_loop_begin": (_i, [_vp, _d, _d, _dp]),
    "azb_loop_begin_sample": (_i, [_vp, _i, _d, _d, _d]),
    "azb_enqueue_steps": (_i, [_vp, _vp, _i]),
    "azb_loop_read": (_i, [_vp, _dp, _ip]),
    "azb_loop_probe": (_i, [_vp]),
    "azb_loop_probe_result": (_i, [_vp, _ip, _dp, _ip]),
    "azb_memcpy_h2d_async": (_i, [_vp, _vp, _sz, _vp]),
    "azb_memcpy_d2h_async": (_i, [_vp, _vp, _sz, _vp]),
    "azb_capture_begin": (_i, [_vp]),
    "azb_capture_end": (_i, [_vp, C.POINTER(_vp)]),
    "azb_graph_launch": (_i, [_vp, _vp]),
    "azb_graph_destroy": (_i, [_vp]),
    "azb_stream_create": (_i, [C.POINTER(_vp)]),
    "azb_stream_destroy": (_i, [_vp]),
    "azb_stream_synchronize": (_i, [_vp]),
    "azb_stream_wait": (_i, [_vp, _vp]),
    "azb_workspace_bytes": (_sz, [_vp]),
    "azb_stream": (_vp, [_vp]),
    "azb_dudt": (_i, [_vp, _vp, _vp]),
    "azb_dudt_noalloc": (_i, [_i, _i, _i, SVParams, _vp, _vp, _vp]),
    "azb_dt": (_i, [_vp, _d, _dp]),
    "azb_u_max": (_i, [_vp, _dp]),
    "azb_rk3_step": (_i, [_vp, _vp, _d, _d, _dp]),
    "azb_simulate_until": (_i, [_vp, _vp, _d, _d, _dp, _i, _i, _ip]),
    "azb_fft_r2c": (_i, [_i, _i, _i, _i, _vp, _vp, _vp]),
    "azb_fft_c2r": (_i, [_i, _i, _i, _i, _vp, _vp, _vp]),
    "azb_fft_work_bytes": (_sz, [_i, _i]),
    "azb_fft_work_init": (_i, [_vp, _i, _i]),
    "azb_fft_r2c_ws": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "azb_fft_c2r_ws": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "azb_copy_to_padded": (_i, [_i, _i, _i, _vp, _vp, _vp]),
    "azb_copy_from_padded": (_i, [_i, _i, _i, _vp, _vp, _vp]),
    "azb_scale": (_i, [_d, _vp, _l, _vp]),
    "azb_copy_bandwidth": (_i, [_d, _i, C.POINTER(_d)]),
    "azb_axpby": (_i, [_d, _vp, _d, _vp, _l, _vp]),
    "azb_max_norm": (_i, [_vp, _l, _dp, _vp]),
    "azb_leray": (_i, [_i, _i, _vp, _vp]),
    "azb_energy_spectrum": (_i, [_i, _i, _vp, _dp, _vp]),
    "azb_malloc": (_i, [C.POINTER(_vp), _sz]),
    "azb_free": (_i, [_vp]),
    "azb_memcpy_h2d": (_i, [_vp, _vp, _sz]),
    "azb_memcpy_d2h": (_i, [_vp, _vp, _sz]),
    "azb_memcpy_d2d": (_i, [_vp, _vp, _sz]),
    "azb_profile_enable": (_i, [_vp, _i]),
    "azb_profile_read": (_i, [_vp, _dp, C.POINTER(C.c_long), _i]),
    "azb_profile_reset": (_i, [_vp]),
    "azb_autotune": (_i, [_vp, _vp, _i]),
    "azb_set_tuning": (_i, [_vp, _ip, _i]),
    "azb_get_tuning": (_i, [_vp, _ip, _ip]),
    "azb_num_plans": (_i, [_vp]),
    "azb_set_plans": (_i, [_vp, _ip]),
    "azb_get_plans": (_i, [_vp, _ip]),
    "azb_autotune_plans": (_i, [_vp, _vp, _i]),
    "azb_synchronize": (_i, [_vp]),
    "azb_device_synchronize": (_i, []),
}

_lib = None


def load() -> C.CDLL:
    """Load libazeban_hip.so once; raise if it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise AzbError(
                f"{LIB_PATH} is not built; run `python -m azeban_amd.build` (hipcc, gfx950)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().azb_last_error()
        raise AzbError(f"azeban_hip error {rc}: {msg.decode() if msg else ''}")
