"""Build and load the oracle's C restatement (oracle/fastvec_ref.c).
TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fastvec_ref.c")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libfastvec_ref.so")


def build() -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["gcc", "-O3", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", SRC, "-o", LIB, "-lm"]
    subprocess.run(cmd, check=True)
    return LIB


def load():
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        build()
    lib = ctypes.CDLL(LIB)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.fastvec_ref.argtypes = [P, I, I, I, D, P, I, D, P, D, P, P]
    lib.fastvec_ref.restype = I
    lib.fastvec_ref_planes.argtypes = [P, I, I, I, D, P, I, D, P, D, P, P, I, I]
    lib.fastvec_ref_planes.restype = I
    return lib


def fastvec_c(lib, src, src_width, src_orig, n, dest_width, dest_orig, dreso, planes=None):
    """(data float64 [n]^3, vec float32 [n]^3 x 3) from the C restatement;
    planes = (x0, x1) computes output planes x0 <= x < x1 only (the others
    stay zero)."""
    s = np.ascontiguousarray(src, dtype=np.float32)
    so = np.ascontiguousarray(np.asarray(src_orig, dtype=np.float32).astype(np.float64))
    do = np.ascontiguousarray(np.asarray(dest_orig, dtype=np.float64))
    data = np.zeros((n, n, n), np.float64)
    vec = np.zeros((n, n, n, 3), np.float32)
    x0, x1 = (0, int(n)) if planes is None else (int(planes[0]), int(planes[1]))
    rc = lib.fastvec_ref_planes(s.ctypes.data, s.shape[0], s.shape[1], s.shape[2], float(np.float32(src_width)),
                                so.ctypes.data, int(n), float(dest_width), do.ctypes.data, float(dreso),
                                data.ctypes.data, vec.ctypes.data, x0, x1)
    assert rc == 0
    return data, vec
