"""Host-side geometry of the search: grid edge, rotation rows and matrices,
refinement neighbourhoods, output translations.  Scalar bookkeeping only;
every voxel-sized computation runs in libvesper_hip.
"""
from __future__ import annotations

import bisect
from functools import lru_cache

import numpy as np
from scipy.spatial.transform import Rotation


@lru_cache(maxsize=None)
def _fast_lengths(limit: int) -> tuple:
    """All 2^a 3^b 5^c 7^d 11^e 13^f <= limit with e + f <= 1, sorted."""
    base = [1]
    for p in (2, 3, 5, 7):
        nxt = []
        for v in base:
            while v <= limit:
                nxt.append(v)
                v *= p
        base = nxt
    out = set(base)
    for v in base:
        for q in (11, 13):
            if v * q <= limit:
                out.add(v * q)
    return tuple(sorted(out))


def next_fast_len(target: int) -> int:
    """pyFFTW 0.12's next_fast_len as used for the grid edge
    (VESPER_1_prob_new.py:97): smallest 2^a 3^b 5^c 7^d 11^e 13^f >= target
    with e + f <= 1."""
    target = max(int(target), 1)
    limit = 64
    while True:
        lens = _fast_lengths(limit)
        i = bisect.bisect_left(lens, target)
        if i < len(lens):
            return lens[i]
        limit *= 4


def calc_angle_comb(ang_interval) -> np.ndarray:
    """Euler rows (x, y, z) in degrees: x, y in [0, 360), z in [0, 180], step
    by repeated addition; z slowest, x middle, y fastest
    (VESPER_1_prob_new.py:1396-1415)."""
    steps = []
    i = 0
    while i < 360:
        steps.append(i)
        i += ang_interval
    zs = []
    i = 0
    while i <= 180:
        zs.append(i)
        i += ang_interval
    # int steps give an integer table, as np.array(meshgrid(...)) does
    xs = np.asarray(steps)
    z = np.asarray(zs)
    nx, nz = len(xs), len(z)
    out = np.empty((nz, nx, nx, 3), dtype=np.result_type(xs, z))
    out[..., 0] = xs[None, :, None]
    out[..., 1] = xs[None, None, :]
    out[..., 2] = z[:, None, None]
    return out.reshape(-1, 3)


def rotation_rows(angles) -> np.ndarray:
    """float32(flip(M)) per Euler row, row-major [R][9], where
    M = R.from_euler("ZYX", angle, degrees=True) with |m| <= 1e-15 set to 0
    (VESPER_1_prob_new.py:305-307); positions use its transpose (:310) and
    vectors use it as is (:341)."""
    a = np.atleast_2d(np.asarray(angles, dtype=np.float64))
    m = Rotation.from_euler("ZYX", a, degrees=True).as_matrix()
    m[np.abs(m) <= 1e-15] = 0.0
    f = m[:, ::-1, ::-1]
    return np.ascontiguousarray(f.reshape(-1, 9), dtype=np.float32)


def refine_rows(angle) -> np.ndarray:
    """+-5 degree neighbourhood of one row in meshgrid order, filtered
    (x < 360, y < 360, z < 180) before negatives wrap by +360
    (VESPER_1_prob_new.py:694-711)."""
    a = np.asarray(angle)
    gx = np.array([a[0] - 5, a[0], a[0] + 5])
    gy = np.array([a[1] - 5, a[1], a[1] + 5])
    gz = np.array([a[2] - 5, a[2], a[2] + 5])
    g = np.empty((3, 3, 3, 3), dtype=np.result_type(gx, gy, gz))
    g[..., 0] = gx[None, :, None]
    g[..., 1] = gy[None, None, :]
    g[..., 2] = gz[:, None, None]
    g = g.reshape(-1, 3)
    g = g[(g[:, 0] < 360) & (g[:, 1] < 360) & (g[:, 2] < 180)]
    g[g < 0] += 360
    return g


def euler2rot(euler):
    """VESPER_1_prob_new.py:1549-1551."""
    return Rotation.from_euler("xyz", euler, degrees=True)


def wrap_trans(trans, dim):
    t = np.array(trans)
    for i in range(3):
        if t[i] > 0.5 * dim:
            t[i] -= dim
    return t


def convert_trans(cen1, cen2, r, trans, xwidth2, dim):
    """Translation in Angstrom for the output lines (VESPER_1_prob_new.py:1554-1566)."""
    t = wrap_trans(trans, dim)
    return cen1 - (r.apply(cen2) + t * xwidth2)
