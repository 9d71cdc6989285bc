"""CPU oracle of the proposal network's arithmetic -- TEST INFRASTRUCTURE ONLY.

float64 forward/backward of the 4-layer MLP with the GPU path's bf16 rounding
points (inputs, weights, hidden activations and hidden-layer gradients are
stored in bf16; every product is accumulated in higher precision), the
mean-squared loss, and the 32 rotational moments of a map.  No reference
counterpart: the learned proposal is an addition of this build.
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round float values to bf16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(X, W, b, w4, b4, round_bf16=True):
    """X [M, IN] (bf16 values); W, b: 3 hidden layers; returns (pred, acts).
    round_bf16=False gives the exact float64 function (finite differences)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    acts = []
    a = r(X)
    for l in range(3):
        z = a @ r(W[l]) + b[l]
        a = r(np.maximum(z, 0.0))
        acts.append(a)
    pred = a @ np.asarray(w4, np.float64) + float(b4[0])
    return pred, acts


def grads(X, y, W, b, w4, b4, round_bf16=True):
    """Gradients (dW1..3, db1..3, dw4, db4) and the loss of mean (pred - y)^2."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    M = X.shape[0]
    pred, acts = forward(X, W, b, w4, b4, round_bf16)
    e = pred - y
    loss = float(np.mean(e * e))
    d = 2.0 * e / M
    dw4 = acts[2].T @ d
    db4 = np.array([d.sum()])
    dz = r(np.where(acts[2] > 0, d[:, None] * np.asarray(w4, np.float64)[None, :], 0.0))
    dW = [None] * 3
    db = [None] * 3
    for l in (2, 1, 0):
        inp = r(X) if l == 0 else acts[l - 1]
        dW[l] = inp.T @ dz
        db[l] = dz.sum(axis=0)
        if l > 0:
            dz = r(np.where(acts[l - 1] > 0, dz @ r(W[l]).T, 0.0))
    return dW + db + [dw4, db4], loss


def moments(data, vec):
    """32 rotational moments of a map about its grid centre n/2 (positions in
    units of n; weights = data > 0)."""
    n = data.shape[0]
    idx = np.argwhere(data > 0)
    d = data[data > 0].astype(np.float64)
    r = (idx - 0.5 * n) / n
    v = vec[data > 0].astype(np.float64)
    w = d.sum() if d.sum() > 0 else 1.0
    c = len(d) if len(d) else 1.0
    out = np.zeros(32)
    out[0:3] = (d[:, None] * r).sum(0) / w
    rr = d[:, None, None] * r[:, :, None] * r[:, None, :]
    m2 = rr.sum(0) / w
    out[3:9] = [m2[0, 0], m2[1, 1], m2[2, 2], m2[0, 1], m2[0, 2], m2[1, 2]]
    out[9:12] = v.sum(0) / c
    out[12:21] = (r[:, :, None] * v[:, None, :]).sum(0).reshape(9) / c
    vv = (v[:, :, None] * v[:, None, :]).sum(0) / c
    out[21:27] = [vv[0, 0], vv[1, 1], vv[2, 2], vv[0, 1], vv[0, 2], vv[1, 2]]
    out[27] = out[3] + out[4] + out[5]
    out[28] = np.linalg.norm(out[0:3])
    out[29] = np.linalg.norm(out[9:12])
    out[30] = out[12] + out[16] + out[20]
    out[31] = len(d) / n ** 3
    return out
