"""CPU oracle of the prediction-grid interpolation -- TEST INFRASTRUCTURE ONLY.

Restates interpolate (interpolation_functions.py:6-135): the reference parity
from the first voxel (C order) whose channel 0 is positive, then seven passes
that fill zero entries of the other parity lattices with float64 averages of
their +-1 neighbours along the differing axes (index -1 wraps around, as
Python/Numba negative indexing does).  `interpolate_loops` follows the
reference's loops element by element (small grids); `interpolate` does each
pass as one vectorised gather, which is valid because a pass reads only
lattices written by earlier passes.
"""
from __future__ import annotations

import numpy as np

# (axes whose parity differs from the reference) in the reference's pass order
PASSES = [(0,), (1,), (2,), (0, 1), (0, 2), (1, 2), (0, 1, 2)]


def reference_parity(data):
    """Parity (x, y, z) of the first voxel in C order with data[..., 0] > 0
    (interpolation_functions.py:8-37); None when there is none."""
    idx = np.flatnonzero(data[..., 0] > 0)
    if idx.size == 0:
        return None
    i, j, k = np.unravel_index(int(idx[0]), data.shape[:3])
    return int(i % 2), int(j % 2), int(k % 2)


def interpolate_loops(data):
    """Element-by-element restatement of the reference's loops (in place)."""
    ref = reference_parity(data)
    if ref is None:
        raise ValueError("no voxel with a positive channel 0")
    dims = data.shape[:3]
    for axes in PASSES:
        starts = [((ref[a] + 1) % 2) if a in axes else ref[a] for a in range(3)]
        for i in range(starts[0], dims[0] - 1, 2):
            for j in range(starts[1], dims[1] - 1, 2):
                for k in range(starts[2], dims[2] - 1, 2):
                    for l in range(data.shape[3]):
                        if data[i, j, k, l] == 0:
                            avg = 0.0
                            p = [i, j, k]
                            for a in axes:
                                q = list(p)
                                q[a] = p[a] + 1
                                avg += float(data[q[0], q[1], q[2], l])
                                q[a] = p[a] - 1          # -1 indexes the last element
                                avg += float(data[q[0], q[1], q[2], l])
                            data[i, j, k, l] = avg / float(2 * len(axes))
    return ref


def interpolate(data):
    """Vectorised form of interpolate_loops (in place); returns the parity."""
    ref = reference_parity(data)
    if ref is None:
        raise ValueError("no voxel with a positive channel 0")
    dims = data.shape[:3]
    for axes in PASSES:
        starts = [((ref[a] + 1) % 2) if a in axes else ref[a] for a in range(3)]
        rng = [np.arange(starts[a], dims[a] - 1, 2) for a in range(3)]
        if any(r.size == 0 for r in rng):
            continue
        I, J, K = np.meshgrid(*rng, indexing="ij")
        cur = data[I, J, K, :]
        avg = np.zeros(cur.shape, dtype=np.float64)
        for a in axes:
            plus = [I, J, K]
            minus = [I, J, K]
            plus[a] = plus[a] + 1
            minus[a] = np.where(minus[a] - 1 < 0, dims[a] - 1, minus[a] - 1)
            avg = avg + data[plus[0], plus[1], plus[2], :].astype(np.float64)
            avg = avg + data[minus[0], minus[1], minus[2], :].astype(np.float64)
        avg = avg / float(2 * len(axes))
        data[I, J, K, :] = np.where(cur == 0, avg.astype(data.dtype), cur)
    return ref
