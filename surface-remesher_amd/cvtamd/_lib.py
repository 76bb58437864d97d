"""Loads lib/libcvt_hip.so (the HIP engine behind include/cvt.h) through ctypes.

There is no fallback: if the library is missing or fails to load, every call
raises, so a run can never silently fall back to CPU code.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_ROOT / "lib" / "libcvt_hip.so"
INCLUDE = PKG_ROOT.parent / "include"

CVT_OK = 0
STATUS = {
    -1: "CVT_ERR_ARG",
    -2: "CVT_ERR_HIP",
    -3: "CVT_ERR_NOMEM",
    -4: "CVT_ERR_STATE",
    -5: "CVT_ERR_CAPACITY",
}
LABEL_NONE = 0x7FFFFFFF


class CvtError(RuntimeError):
    def __init__(self, status: int, func: str, msg: str):
        super().__init__(f"{func}: {STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None

_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_d = C.c_double

# name -> (restype, argtypes); every entry point declared in include/*.h
SIGNATURES = {
    "cvt_last_error": (C.c_char_p, []),
    "cvt_version": (_i, []),
    "cvt_create": (_i, [C.POINTER(_vp), _i, _i, _i]),
    "cvt_destroy": (_i, [_vp]),
    "cvt_set_stream": (_i, [_vp, _vp]),
    "cvt_set_density": (_i, [_vp, _vp]),
    "cvt_set_sites": (_i, [_vp, _vp, _vp, _i]),
    "cvt_get_sites": (_i, [_vp, _vp]),
    "cvt_num_sites": (_i, [_vp]),
    "cvt_voronoi": (_i, [_vp]),
    "cvt_get_labels": (_i, [_vp, _vp]),
    "cvt_moments": (_i, [_vp, _vp]),
    "cvt_lloyd": (_i, [_vp, _i, _f]),
    "cvt_lloyd_until": (_i, [_vp, _f, _d, _i, _d, _i, C.POINTER(_i)]),
    "cvt_history_length": (_i, [_vp]),
    "cvt_get_history": (_i, [_vp, _vp, _i]),
    "cvt_clear_history": (_i, [_vp]),
    "cvt_energy": (_i, [_vp, C.POINTER(_d), C.POINTER(_d)]),
    "cvt_synchronize": (_i, [_vp]),
    "cvt_profile": (_i, [_vp, _i, _f, _vp]),
    "cvt_compat_gcvt": (_i, [_vp, _vp, _vp, _i, _i, _i, C.POINTER(_i), C.POINTER(_i)]),
    "cvt_ref_init_sites": (_i, [_vp, _vp, _vp, _i, _i, _vp, C.POINTER(C.c_longlong)]),
    "cvt_rasterize": (_i, [_vp, _vp, _i, _vp, _i, _vp, _d, _i]),
    "cvt_rasterize_noalloc": (_i, [_vp, _vp, _i, _vp, _i, _vp, _d, _i]),
    "cvt_extract_points": (_i, [_vp, _vp, _i, _d, _d, _d, _vp, C.c_long, C.POINTER(C.c_long)]),
    "cvt_locate_points": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, _vp, _vp]),
    "cvt_vertex_weights": (_i, [_vp, _vp, _i, _vp, _i, _vp]),
    "cvt_batch_create": (_i, [C.POINTER(_vp), _i, _vp, _vp, _vp]),
    "cvt_batch_destroy": (_i, [_vp]),
    "cvt_batch_set_stream": (_i, [_vp, _vp]),
    "cvt_batch_set_density": (_i, [_vp, _i, _vp]),
    "cvt_batch_set_sites": (_i, [_vp, _i, _vp, _vp, _i]),
    "cvt_batch_get_sites": (_i, [_vp, _i, _vp]),
    "cvt_batch_lloyd": (_i, [_vp, _i, _f]),
    "cvt_batch_voronoi": (_i, [_vp]),
    "cvt_batch_get_labels": (_i, [_vp, _i, _vp]),
    "cvt_batch_history": (_i, [_vp, _i, _vp, _i]),
    "cvt_batch_set_tile": (_i, [_vp, _i, _i]),
    "cvt_batch_get_tile": (_i, [_vp, _i]),
    "cvt_batch_synchronize": (_i, [_vp]),
    "cvt_unet_create": (_i, [C.POINTER(_vp), _i, _i, C.c_ulonglong]),
    "cvt_unet_destroy": (_i, [_vp]),
    "cvt_unet_set_stream": (_i, [_vp, _vp]),
    "cvt_unet_num_params": (C.c_long, [_vp]),
    "cvt_unet_get_params": (_i, [_vp, _vp]),
    "cvt_unet_set_params": (_i, [_vp, _vp]),
    "cvt_unet_get_grads": (_i, [_vp, _vp]),
    "cvt_unet_prepare": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp]),
    "cvt_unet_set_input": (_i, [_vp, _vp, _i]),
    "cvt_unet_forward": (_i, [_vp, _i]),
    "cvt_unet_get_field": (_i, [_vp, _vp, _i]),
    "cvt_unet_loss_backward": (_i, [_vp, _i, _vp, _vp, _i, C.POINTER(_f)]),
    "cvt_unet_adam": (_i, [_vp, _f, _f, _f, _f]),
    "cvt_unet_sample": (_i, [_vp, _i, _vp, _i, _vp]),
    "cvt_unet_warm_start": (_i, [_vp, _vp]),
    "cvt_unet_save": (_i, [_vp, C.c_char_p]),
    "cvt_unet_load": (_i, [_vp, C.c_char_p]),
    "cvt_unet_synchronize": (_i, [_vp]),
    "cvt_set_tile": (_i, [_vp, _i]),
    "cvt_get_tile": (_i, [_vp]),
    "cvt_dt_create": (_i, [C.POINTER(_vp), _i]),
    "cvt_dt_compute": (_i, [_vp, _vp, _i, _vp, _vp, _i, C.POINTER(_i)]),
    "cvt_dt_compute_constrained": (_i, [_vp, _vp, _i, _vp, _i, _vp, _vp, _i, C.POINTER(_i)]),
    "cvt_dt_stats": (_i, [_vp, _vp]),
    "cvt_dt_destroy": (_i, [_vp]),
    "cvt_dt_predicates": (_i, [_vp, _i, _vp, _vp]),
    "cvt_set_jfa_shape": (_i, [_vp, _i, _i]),
    "cvt_get_jfa_shape": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
}


def lib():
    """The loaded engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
            f"or make -C {PKG_ROOT})")
    try:
        # torch ships its own HIP runtime; loaded after ours, it fails to
        # initialise the device, so it goes first when it is installed
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status: int, func: str) -> int:
    if status < 0:
        msg = lib().cvt_last_error().decode(errors="replace")
        raise CvtError(status, func, msg)
    return status


def call(name: str, *args) -> int:
    return check(getattr(lib(), name)(*args), name)
