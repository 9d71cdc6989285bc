"""The C-ABI library loads on the host and exports every entry point that
include/vesper_hip.h declares (no GPU compute here)."""
import ctypes
import os
import re

from vesper_python_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vesper_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vx_[a-z0-9_]+)\s*\(", src)))


def _lib():
    if not os.path.exists(_native.LIB_PATH):
        from vesper_python_amd import build
        build.build(verbose=False)
    return _native.load()


def test_header_declares_the_hot_path():
    names = _declared()
    for must in ("vx_fastvec", "vx_score_rows", "vx_topn", "vx_rotate_map", "vx_get_score",
                 "vx_set_target", "vx_set_search", "vx_plan_create", "vx_map_stats"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib()
    raw = ctypes.CDLL(_native.LIB_PATH)
    for name in _declared():
        assert hasattr(raw, name), name
        assert name in _native.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.vx_abi_version() == _native.ABI_VERSION


def test_errors_are_status_codes_not_aborts():
    lib = _lib()
    # argument validation happens before any device call
    h = ctypes.c_void_p()
    rc = lib.vx_plan_create(0, 0, 1, 0, ctypes.byref(h))
    assert rc == -1 and b"grid edge" in lib.vx_last_error()
    rc = lib.vx_plan_create(128, 99, 1, 0, ctypes.byref(h))
    assert rc == -1 and b"mode" in lib.vx_last_error()
    rc = lib.vx_topn(None, 10, 1, 0, 5, None, None)
    assert rc == -1
    rc = lib.vx_query_planes(None, None, None, None)
    assert rc == -1 and b"NULL" in lib.vx_last_error()
    assert lib.vx_plan_destroy(None) == 0


def test_topn_workspace_size_is_host_only():
    lib = _lib()
    nb = ctypes.c_int64(0)
    assert lib.vx_topn_ws_bytes(24624, 10, ctypes.byref(nb)) == 0
    # two buffers of (13 chunks x 10 keys + 2048) 8-byte keys
    assert nb.value == 2 * (13 * 10 + 2048) * 8
    assert lib.vx_topn_ws_bytes(0, 10, ctypes.byref(nb)) == -1
    assert lib.vx_topn_ws_bytes(10, 1025, ctypes.byref(nb)) == -1
