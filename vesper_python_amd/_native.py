"""ctypes binding of libvesper_hip.so (declared in include/vesper_hip.h).

The library is the only compute path: if it is missing, or no GPU is
visible, every call raises VesperError -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libvesper_hip.so")

VX_MODE_VEC, VX_MODE_OVERLAP, VX_MODE_CC, VX_MODE_PCC, VX_MODE_LAPLACIAN, VX_MODE_PROB = range(6)
VX_FLAG_NO_FMA = 1
VX_FLAG_NO_PRUNE = 2
VX_TUNE_XPASS_VARIANT = 1
VX_TUNE_XPASS_TILE = 2
VX_TUNE_FORCE_PENCIL = 3
VX_TUNE_PIN_BOX_WIDTH = 4
VX_TUNE_PIN_BOX_LO = 5
VX_TUNE_SLAB_THREADS = 6
VX_TUNE_STREAMS = 7
VX_TUNE_STATIC_PLANS = 8
VX_TUNE_RADIX_ORDER = 9
VX_TUNE_SLAB_FUSED = 10
VX_TUNE_PENCIL_CHUNKS = 11
ABI_VERSION = 1

MODE_CODES = {"VecProduct": VX_MODE_VEC, "Overlap": VX_MODE_OVERLAP, "CC": VX_MODE_CC,
              "PCC": VX_MODE_PCC, "Laplacian": VX_MODE_LAPLACIAN}

# name -> (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_int64
SIGNATURES = {
    "vx_abi_version": (_I, []),
    "vx_last_error": (ctypes.c_char_p, []),
    "vx_fastvec": (_I, [_P, _I, _I, _I, _D, _P, _I, _D, _P, _D, _P, _P, _P]),
    "vx_map_stats": (_I, [_P, _L, _P, _P]),
    "vx_plan_create": (_I, [_I, _I, _I, _I, ctypes.POINTER(_P)]),
    "vx_plan_destroy": (_I, [_P]),
    "vx_plan_info": (_I, [_P, _P]),
    "vx_plan_tune": (_I, [_P, _I, _I]),
    "vx_set_target": (_I, [_P, _P, _P, _P, _D, _P]),
    "vx_set_search": (_I, [_P, _P, _P, _P, _D, _P]),
    "vx_set_prob_stats": (_I, [_P, _I, _D, _D, _D, _D, _D]),
    "vx_score_rows": (_I, [_P, _P, _I, _P, _P, _P]),
    "vx_query_planes": (_I, [_P, _P, _P, _P]),
    "vx_score_rows_profile": (_I, [_P, _P, _I, _P, _P, _P, _P]),
    "vx_rescore_f64": (_I, [_P, _P, _P, _P, _D, _P, _P, _P, _P, _I, _D, _I, _P, _P, _P, _P]),
    "vx_topn": (_I, [_P, _I, _I, _I, _I, _P, _P]),
    "vx_topn_ws_bytes": (_I, [_I, _I, _P]),
    "vx_topn_ws": (_I, [_P, _I, _I, _I, _I, _P, _P, ctypes.c_int64, _P]),
    "vx_topn_noalloc": (_I, [_P, _I, _I, _I, _I, _P, _P]),
    "vx_rotate_map": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _I, _P]),
    "vx_get_score": (_I, [_I, _P, _P, _D, _P, _P, _P, _P, _P]),
    "vx_laplacian": (_I, [_I, _P, _P, _P]),
    "vx_mode_planes": (_I, [_I, _I, _P, _P, _P, _D, _P, _P]),
    "vx_interpolate": (_I, [_P, _I, _I, _I, _I, _I, _P, _P]),
    "vx_correlate": (_I, [_P, _P, _I, _P, _P, _P]),
    "vx_sum_argmax": (_I, [_P, _I, ctypes.c_int64, _P, _P, _P]),
    "vx_mixed_argmax": (_I, [_P, _I, _P, _I, ctypes.c_int64, _D, _D, _D, _D, _D, _P, _P, _P]),
    "vx_sco_arr": (_I, [_I, _P, _P, _P, _I, _P, _P]),
    "vx_threshold_dmax": (_I, [_P, _I, _I, _I, ctypes.c_float, _P, _P]),
    "vx_gemm_bf16": (_I, [_I, _I, _I, _P, _I, _I, _P, _I, _I, _P, _I, _I, _P, _P, _I, _P]),
    "vx_mlp_head_fwd": (_I, [_I, _I, _P, _I, _P, _P, _P, _P]),
    "vx_mlp_head_bwd": (_I, [_I, _I, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "vx_colsum_bf16": (_I, [_I, _I, _P, _I, _P, _P]),
    "vx_adam": (_I, [_I, _P, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                     _I, ctypes.c_float, _P]),
    "vx_cast_bf16": (_I, [_I, _P, _P, _P]),
    "vx_mlp_descriptors": (_I, [_I, _P, _P, _P, _P]),
    "vx_mlp_features": (_I, [_I, _I, _P, _P, _P, _P, _P, _P]),
    "vx_select_top": (_I, [_I, _P, _I, _P, _P, _P]),
    "vx_zscore_f32": (_I, [_I, _P, _P, _I, _P, _P]),
    "vx_gather_rows": (_I, [_I, _I, _P, _P, _I, _P, _P]),
    "vx_mask_scored": (_I, [_I, _P, _P, _P]),
    "vx_filter_unscored": (_I, [_I, _P, _P, _P, _P, _P]),
    "vx_scatter_scores": (_I, [_I, _P, _P, _I, _P, _P, _P]),
    "vx_fill_f32": (_I, [_I, ctypes.c_float, _P, _P]),
}


class VesperError(RuntimeError):
    """A failure reported by libvesper_hip (or its absence)."""


_lock = threading.Lock()
_lib = None


def load(path: str = LIB_PATH):
    """Load the library once; raise VesperError if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise VesperError(f"{path} is not built: run `python -m vesper_python_amd.build` "
                              "(this package has no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.vx_abi_version() != ABI_VERSION:
            raise VesperError(f"ABI mismatch: library {lib.vx_abi_version()} != {ABI_VERSION}")
        _lib = lib
        return lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().vx_last_error().decode(errors="replace")
        raise VesperError(f"{what} failed ({rc}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
