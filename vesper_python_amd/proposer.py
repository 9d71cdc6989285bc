"""Learned rotation proposal: which Euler rows deserve a full FFT score.

A 4-layer MLP (80 -> 256 -> 256 -> 256 -> 1; bf16 weights and activations,
fp32 accumulation on MFMA through vx_gemm_bf16) maps a rotation's quaternion
plus 64 rotational moments (32 of the target field, 32 of the query field
rotated by that rotation) to a predicted DOT score.  It is trained with the
library's own forward/backward kernels and Adam on score landscapes the
exhaustive search produces: across training pairs, and per query pair on a
coarse landscape (every row of a coarser angle step) before proposing rows of
the fine step.  Only the predicted top fraction of fine rows is FFT-scored;
the exhaustive search stays available as the fallback and as the
reference its top-10 recall is measured against.  No reference counterpart: the reference scores every row
(VESPER_1_prob_new.py:644-655).
"""
from __future__ import annotations

import math

import numpy as np
import torch
from scipy.spatial.transform import Rotation

from . import _native as nat
from . import engine
from .geometry import calc_angle_comb, rotation_rows

IN = 80            # 4 + 32 + 32 features, padded to a multiple of 16
H = 256


def _p(t):
    return engine._p(t)


def _st():
    return engine._stream()


def pad64(m: int) -> int:
    return (m + 63) // 64 * 64


def quaternions(angles) -> np.ndarray:
    """Unit quaternion (x, y, z, w), w >= 0, of the content rotation
    R_xyz(angle) that a row applies to the query (euler2rot,
    VESPER_1_prob_new.py:1549-1551)."""
    q = Rotation.from_euler("xyz", np.atleast_2d(angles), degrees=True).as_quat()
    q[q[:, 3] < 0] *= -1
    return np.ascontiguousarray(q, dtype=np.float32)


def descriptors(data: torch.Tensor, vec: torch.Tensor) -> torch.Tensor:
    """32 rotational moments of a resampled map (device float32 [32])."""
    out = torch.empty(32, dtype=torch.float32, device=data.device)
    nat.call("vx_mlp_descriptors", int(data.shape[0]), _p(data), _p(vec), _p(out), _st())
    return out


def features(rot: torch.Tensor, quat: torch.Tensor, tdesc: torch.Tensor, qdesc: torch.Tensor,
             out: torch.Tensor | None = None) -> torch.Tensor:
    """bf16 feature rows [pad64(R), IN] (zero rows past R), or written into
    the first R rows of `out`."""
    R = int(rot.shape[0])
    X = torch.zeros((pad64(R), IN), dtype=torch.bfloat16, device=rot.device) if out is None else out
    nat.call("vx_mlp_features", R, IN, _p(rot), _p(quat), _p(tdesc), _p(qdesc), _p(X), _st())
    return X


def gather_rows(src: torch.Tensor, idx: torch.Tensor, n: int, out: torch.Tensor, nout: int | None = None):
    """out rows [0, n) = src rows idx[:n] (idx int32 on the device), zero rows
    up to nout (vx_gather_rows)."""
    row_bytes = src.element_size() * (src[0].numel() if src.dim() > 1 else 1)
    nat.call("vx_gather_rows", int(n if nout is None else nout), int(n), _p(idx), _p(src), int(row_bytes), _p(out),
             _st())
    return out


def zscore_into(s: torch.Tensor, n: int, out: torch.Tensor, stats: torch.Tensor, given: bool = False):
    """out[:n] = (s[:n] - mean) / (std + 1e-30) (vx_zscore_f32); stats
    (float64 [2] on the device) receives mean and std + 1e-30, or, given,
    supplies them."""
    nat.call("vx_zscore_f32", int(n), _p(s), _p(stats), int(bool(given)), _p(out), _st())
    return out


class RotationMLP:
    """The proposal network; parameters are fp32 masters with bf16 copies."""

    def __init__(self, seed: int = 0, device=None):
        dev = device or engine.require_gpu()
        rng = np.random.default_rng(seed)
        fan = [IN, H, H]
        self.W = [torch.from_numpy((rng.standard_normal((k, H)) * math.sqrt(2.0 / k)).astype(np.float32)).to(dev)
                  for k in fan]
        self.b = [torch.zeros(H, dtype=torch.float32, device=dev) for _ in fan]
        self.w4 = torch.from_numpy((rng.standard_normal(H) * math.sqrt(1.0 / H)).astype(np.float32)).to(dev)
        self.b4 = torch.zeros(1, dtype=torch.float32, device=dev)
        self.Wb = [torch.empty((k, H), dtype=torch.bfloat16, device=dev) for k in fan]
        for w, wb in zip(self.W, self.Wb):
            nat.call("vx_cast_bf16", w.numel(), _p(w), _p(wb), _st())
        self.params = self.W + self.b + [self.w4, self.b4]
        self.m1 = [torch.zeros_like(p) for p in self.params]
        self.m2 = [torch.zeros_like(p) for p in self.params]
        self.steps = 0
        self.dev = dev
        self._bufs = {}

    # -- buffers ---------------------------------------------------------------
    def _get(self, Mp: int):
        if Mp not in self._bufs:
            z = lambda: torch.zeros((Mp, H), dtype=torch.bfloat16, device=self.dev)
            self._bufs[Mp] = dict(A=[z(), z(), z()], D=[z(), z(), z()],
                                  pred=torch.zeros(Mp, dtype=torch.float32, device=self.dev))
        return self._bufs[Mp]

    # -- forward ---------------------------------------------------------------
    def forward(self, X: torch.Tensor, M: int) -> torch.Tensor:
        """Predicted scores of the first M rows of X (bf16 [Mp, IN])."""
        Mp = int(X.shape[0])
        buf = self._get(Mp)
        inp, ldi = X, IN
        for l in range(3):
            K = IN if l == 0 else H
            nat.call("vx_gemm_bf16", Mp, H, K, _p(inp), ldi, 0, _p(self.Wb[l]), H, 0, _p(buf["A"][l]), H, 1,
                     _p(self.b[l]), None, 0, _st())
            inp, ldi = buf["A"][l], H
        nat.call("vx_mlp_head_fwd", M, H, _p(buf["A"][2]), H, _p(self.w4), _p(self.b4), _p(buf["pred"]), _st())
        return buf["pred"][:M]

    # -- training --------------------------------------------------------------
    def grads(self, X: torch.Tensor, y: torch.Tensor, M: int, loss: torch.Tensor | None = None):
        """Mean-squared-loss gradients of every parameter (fp32, same order as
        self.params) and the loss (added to `loss`, a device float32 [1],
        when given), by the library's backward kernels."""
        Mp = int(X.shape[0])
        buf = self._get(Mp)
        pred = self.forward(X, M)
        g = [torch.zeros_like(p) for p in self.params]
        gW, gb, gw4, gb4 = g[0:3], g[3:6], g[6], g[7]
        if loss is None:
            loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        A, D = buf["A"], buf["D"]
        if Mp > M:            # rows past M take no part (the head writes rows < M)
            nat.call("vx_gather_rows", Mp - M, 0, None, None, H * 2, _p(D[2][M:]), _st())
        # output layer; D[2] = dL/dZ3 (through the ReLU of A3)
        nat.call("vx_mlp_head_bwd", M, H, _p(A[2]), H, _p(self.w4), _p(pred), _p(y), _p(gw4), _p(gb4),
                 _p(D[2]), _p(loss), _st())
        for l in (2, 1, 0):
            inp, K = (X, IN) if l == 0 else (A[l - 1], H)
            # dW_l += inp^T dZ_l
            nat.call("vx_gemm_bf16", K, H, Mp, _p(inp), K, 1, _p(D[l]), H, 0, _p(gW[l]), H, 3, None, None, 0, _st())
            nat.call("vx_colsum_bf16", Mp, H, _p(D[l]), H, _p(gb[l]), _st())
            if l > 0:
                # dZ_{l-1} = (dZ_l W_l^T) * (A_{l-1} > 0)
                nat.call("vx_gemm_bf16", Mp, H, H, _p(D[l]), H, 0, _p(self.Wb[l]), H, 1, _p(D[l - 1]), H, 2,
                         None, _p(A[l - 1]), H, _st())
        return g, loss

    def train_step(self, X, y, M, lr=1e-3, wd=0.0, loss=None):
        g, loss = self.grads(X, y, M, loss)
        self.steps += 1
        bf = self.Wb + [None] * 5
        for p, gp, a, b, pb in zip(self.params, g, self.m1, self.m2, bf):
            nat.call("vx_adam", p.numel(), _p(p), _p(gp), _p(a), _p(b), _p(pb), lr, 0.9, 0.999, 1e-8,
                     self.steps, wd, _st())
        return loss

    def fit(self, X, y, M, epochs=20, batch=2048, lr=1e-3, seed=0):
        """Minibatch Adam over the rows of X (bf16 [>= M, IN]) / y (fp32
        [>= M]); minibatches gathered on the device (vx_gather_rows).
        Returns the mean minibatch loss of each epoch."""
        rng = np.random.default_rng(seed)
        bp = pad64(min(batch, M))
        xb = torch.zeros((bp, IN), dtype=torch.bfloat16, device=self.dev)
        yb = torch.zeros(bp, dtype=torch.float32, device=self.dev)
        # per-epoch loss sums stay on the device until the end (no host sync
        # per epoch)
        tot = torch.zeros(epochs, device=self.dev)
        nb = (M + batch - 1) // batch
        for e in range(epochs):
            perm = torch.from_numpy(rng.permutation(M).astype(np.int32)).to(self.dev)
            for s in range(0, M, batch):
                m = min(batch, M - s)
                idx = perm[s:s + m]
                gather_rows(X, idx, m, xb, bp)
                gather_rows(y, idx, m, yb, bp)
                self.train_step(xb, yb, m, lr, loss=tot[e:e + 1])
        return [float(v) / nb for v in tot.cpu().numpy()]

    def state(self) -> dict:
        return {f"p{i}": p.cpu().numpy() for i, p in enumerate(self.params)}

    def load(self, st: dict):
        """Parameters from `st`; the optimiser starts afresh (Adam moments
        zero, step count 0), so a fit after load does not depend on what the
        network was fitted to before."""
        for i, p in enumerate(self.params):
            p.copy_(torch.from_numpy(st[f"p{i}"]))
        for w, wb in zip(self.W, self.Wb):
            nat.call("vx_cast_bf16", w.numel(), _p(w), _p(wb), _st())
        for a, b in zip(self.m1, self.m2):
            a.zero_()
            b.zero_()
        self.steps = 0


def landscape(plan: engine.SearchPlan, rot: torch.Tensor) -> torch.Tensor:
    """Exhaustive float32 best scores of every row (the training labels)."""
    best, _ = plan.score_rows(rot)
    return best.view(-1) if best.shape[1] == 1 else best[:, 0].contiguous()


def zscore(s: torch.Tensor) -> torch.Tensor:
    """(s - mean) / (std + 1e-30) on the device."""
    out = torch.empty_like(s)
    return zscore_into(s, int(s.numel()), out, torch.empty(2, dtype=torch.float64, device=s.device))


class Proposer:
    """Proposal search for one pair held in a search plan."""

    def __init__(self, plan: engine.SearchPlan, tdata, tvec, sdata, svec, mlp: RotationMLP):
        self.plan = plan
        self.mlp = mlp
        self.td = descriptors(tdata, tvec)
        self.qd = descriptors(sdata, svec)

    def rows_features(self, angles, out=None):
        rot = torch.from_numpy(rotation_rows(angles)).to(self.td.device)
        quat = torch.from_numpy(quaternions(angles)).to(self.td.device)
        return rot, features(rot, quat, self.td, self.qd, out)

    def fine_tune(self, coarse_angles, epochs=30, lr=1e-3):
        """Train on this pair's own coarse landscape (every row of a coarse
        angle step, FFT-scored by the plan)."""
        rot, X = self.rows_features(coarse_angles)
        y = zscore(landscape(self.plan, rot))
        return self.mlp.fit(X, y, int(rot.shape[0]), epochs=epochs, batch=512, lr=lr)

    def propose(self, angles, frac=0.05, topn=10):
        """FFT-score only the predicted top `frac` of `angles`; returns
        (top rows into angles, their float32 best, number of rows scored)."""
        rot, X = self.rows_features(angles)
        R = int(rot.shape[0])
        pred = self.mlp.forward(X, R)
        k = max(topn, int(math.ceil(frac * R)))
        rows = torch.empty(R, dtype=torch.int32, device=rot.device)
        nsel = np.zeros(1, dtype=np.int32)
        nat.call("vx_select_top", R, _p(pred), k, _p(rows), nsel.ctypes.data_as(__import__("ctypes").c_void_p), _st())
        ns = int(nsel[0])
        rsel = gather_rows(rot, rows, ns, torch.empty((ns, 9), dtype=torch.float32, device=rot.device))
        best, _ = self.plan.score_rows(rsel)
        b = best.view(-1)
        t = engine.topn(b, topn)
        return rows[:ns].cpu().numpy()[t], b, ns


def pretrain(mlp: RotationMLP, pairs, angles, epochs=20, lr=2e-3, batch=2048, seed=0):
    """Train `mlp` across pairs: every row of `angles` of each pair (a
    Proposer holding the pair's plan and maps) is FFT-scored exhaustively and
    z-scored within its pair; the rows of all pairs form one training set.
    Returns the per-epoch losses."""
    R = len(angles)
    M = R * len(pairs)
    dev = mlp.dev
    Xp = torch.zeros((pad64(M), IN), dtype=torch.bfloat16, device=dev)
    y = torch.empty(M, dtype=torch.float32, device=dev)
    stats = torch.empty(2, dtype=torch.float64, device=dev)
    for k, pr in enumerate(pairs):
        rot, _ = pr.rows_features(angles, out=Xp[k * R:])
        zscore_into(landscape(pr.plan, rot), R, y[k * R:], stats)
    return mlp.fit(Xp, y, M, epochs=epochs, batch=batch, lr=lr, seed=seed)


def pretrain_streamed(mlp: RotationMLP, make_proposer, n_pairs: int, angles, epochs=20, lr=2e-3, batch=2048,
                      seed=0, progress=None):
    """pretrain over n_pairs pairs made one at a time: make_proposer(k)
    returns the k-th pair's Proposer, whose features and z-scored exhaustive
    landscape go into the training set before its plan is closed, so that one
    plan at a time holds device memory (the 512-pair training set of the
    protocol is 12.6 M rows).  progress(k, grid_edge) after each pair.
    Returns the per-epoch losses."""
    R = len(angles)
    M = R * n_pairs
    dev = mlp.dev
    Xp = torch.zeros((pad64(M), IN), dtype=torch.bfloat16, device=dev)
    y = torch.empty(M, dtype=torch.float32, device=dev)
    stats = torch.empty(2, dtype=torch.float64, device=dev)
    for k in range(n_pairs):
        pr = make_proposer(k)
        rot, _ = pr.rows_features(angles, out=Xp[k * R:])
        zscore_into(landscape(pr.plan, rot), R, y[k * R:], stats)
        n = pr.plan.n
        pr.plan.close()
        if progress:
            progress(k, n)
    return mlp.fit(Xp, y, M, epochs=epochs, batch=batch, lr=lr, seed=seed)


def grid_shape(angles):
    """(nx, nz) when `angles` is a calc_angle_comb grid (row = iz nx^2 + ix nx + iy
    for angle (x[ix], x[iy], z[iz])), else None."""
    a = np.asarray(angles, dtype=np.float64)
    xs = np.unique(a[:, 0])
    zs = np.unique(a[:, 2])
    nx, nz = len(xs), len(zs)
    if nx < 2 or len(a) != nx * nx * nz:
        return None
    iz, rem = np.divmod(np.arange(len(a)), nx * nx)
    ix, iy = np.divmod(rem, nx)
    if not (np.array_equal(a[:, 0], xs[ix]) and np.array_equal(a[:, 1], xs[iy]) and np.array_equal(a[:, 2], zs[iz])):
        return None
    return nx, nz


def neighbour_rows(rows, nx, nz):
    """Grid neighbours (+-1 step per Euler angle; x and y wrap at 360, z is
    clamped to the grid) of the given rows of a calc_angle_comb grid."""
    rows = np.asarray(rows, dtype=np.int64)
    iz, rem = np.divmod(rows, nx * nx)
    ix, iy = np.divmod(rem, nx)
    out = []
    for dz in (-1, 0, 1):
        z = iz + dz
        ok = (z >= 0) & (z < nz)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                r = z * nx * nx + ((ix + dx) % nx) * nx + (iy + dy) % nx
                out.append(r[ok])
    return np.unique(np.concatenate(out))


def nearest_coarse_rows(angles, coarse_step):
    """For each Euler row, the calc_angle_comb row of the coarse grid (step
    `coarse_step`) nearest to it (x and y wrap at 360, z is clamped to
    [0, 180]), and its distance in degrees."""
    a = np.asarray(angles, dtype=np.float64)
    nx = len(np.arange(0, 360, coarse_step))
    nz = len(np.arange(0, 180 + 1e-9, coarse_step))
    ix = np.rint(a[:, 0] / coarse_step).astype(np.int64) % nx
    iy = np.rint(a[:, 1] / coarse_step).astype(np.int64) % nx
    iz = np.clip(np.rint(a[:, 2] / coarse_step).astype(np.int64), 0, nz - 1)
    d = np.stack([a[:, 0] - ix * coarse_step, a[:, 1] - iy * coarse_step, a[:, 2] - iz * coarse_step], axis=1)
    d = (d + 180.0) % 360.0 - 180.0
    return iz * nx * nx + ix * nx + iy, np.sqrt((d * d).sum(axis=1))


POLICIES = ("mlp", "coarse", "random", "hybrid")


def propose_rounds(pr: "Proposer", coarse_angles, angles, frac=0.05, rounds=2, topn=10, epochs=40, lr=2e-3,
                   seed=0, sweeps=3, policy="mlp", coarse_step=None):
    """Coarse landscape -> train -> propose a share of the rows -> FFT-score
    them -> retrain on everything scored so far -> propose the next share of
    the rows not yet scored ... (`rounds` shares of frac / rounds each); then
    up to `sweeps` passes that FFT-score the unscored grid neighbours of the
    current top rows (hill climbing on the exact scores).
    `policy` picks the shares: "mlp" (the network's predictions), or, for the
    ablation with the same budget and sweeps, "coarse" (each fine row ranked
    by the coarse landscape's score of its nearest coarse row, nearer rows
    first; coarse_angles must then be the calc_angle_comb grid of step
    `coarse_step`) or "random" (rows drawn uniformly); "hybrid" spends half of
    each share on the coarse ranking (scored first, so the network retrains
    on them) and the other half on the network's predictions: the coarse
    landscape's best basins are always explored, and the network adds the
    rows it ranks high elsewhere.
    Returns (top rows into angles, number of fine rows FFT-scored, coarse rows)."""
    if policy not in POLICIES:
        raise ValueError(f"unknown policy {policy!r}")
    dev = pr.td.device
    Rc = len(coarse_angles)
    R = len(angles)
    # training set: the coarse rows, then every fine row scored (capacity)
    trainX = torch.zeros((pad64(Rc + R), IN), dtype=torch.bfloat16, device=dev)
    trainy = torch.empty(Rc + R, dtype=torch.float32, device=dev)
    cstats = torch.empty(2, dtype=torch.float64, device=dev)
    crot, _ = pr.rows_features(coarse_angles, out=trainX)
    cs = landscape(pr.plan, crot)
    zscore_into(cs, Rc, trainy, cstats)
    ntrain = Rc
    rot, X = pr.rows_features(angles)
    scored = torch.zeros(R, dtype=torch.uint8, device=dev)
    best_all = torch.empty(R, dtype=torch.float32, device=dev)
    nat.call("vx_fill_f32", R, -float("inf"), _p(best_all), _st())
    rows = torch.empty(R, dtype=torch.int32, device=dev)
    sel = torch.empty(R, dtype=torch.int32, device=dev)
    rsel = torch.empty((R, 9), dtype=torch.float32, device=dev)
    pred_buf = torch.empty(R, dtype=torch.float32, device=dev)
    nscored = 0
    share = max(topn, int(math.ceil(frac * R / rounds)))
    static_pred = None
    coarse_pred = None
    if policy in ("coarse", "hybrid"):
        near, dist = nearest_coarse_rows(angles, coarse_step)
        base = trainy[:Rc].cpu().numpy().astype(np.float64)[near]
        coarse_pred = torch.from_numpy((base - 1e-6 * dist).astype(np.float32)).to(dev)
        if policy == "coarse":
            static_pred = coarse_pred
    elif policy == "random":
        static_pred = torch.from_numpy(np.random.default_rng(seed).random(R).astype(np.float32)).to(dev)
    cnt = np.zeros(1, dtype=np.int32)
    cnt_p = cnt.ctypes.data_as(__import__("ctypes").c_void_p)

    def score_new(n):
        """FFT-score the n rows in sel; record them and their scores."""
        gather_rows(rot, sel, n, rsel)
        b, _ = pr.plan.score_rows(rsel[:n])
        nat.call("vx_scatter_scores", n, _p(sel), _p(b), int(b.shape[1]), _p(best_all), _p(scored), _st())
        return b

    def take(pred, k):
        """FFT-score the unscored rows ranked highest by pred (in place: rows
        already scored drop just below the lowest unscored prediction) -- the
        top k and the rest of their histogram bin -- and add them to the
        training set; returns how many were scored."""
        nonlocal nscored, ntrain
        nat.call("vx_mask_scored", R, _p(pred), _p(scored), _st())
        nat.call("vx_select_top", R, _p(pred), k, _p(rows), cnt_p, _st())
        nat.call("vx_filter_unscored", int(cnt[0]), _p(rows), _p(scored), _p(sel), cnt_p, _st())
        n = int(cnt[0])
        if n == 0:
            return 0
        b = score_new(n)
        nscored += n
        gather_rows(X, sel, n, trainX[ntrain:])
        # labels on the coarse landscape's scale
        zscore_into(b, n, trainy[ntrain:], cstats, given=True)
        ntrain += n
        return n

    for r in range(rounds):
        taken = 0
        if policy == "hybrid":
            # the coarse half (its threshold bin may overshoot; the network
            # gets what is left of the share, possibly nothing)
            taken = take(pred_buf.copy_(coarse_pred), (share + 1) // 2)
        if static_pred is None:
            pr.mlp.fit(trainX, trainy, ntrain, epochs=epochs if r == 0 else max(5, epochs // 4), batch=512, lr=lr,
                       seed=seed + r)
            pred = pr.mlp.forward(X, R)
        else:
            pred = pred_buf.copy_(static_pred)
        if share - taken > 0:
            taken += take(pred, share - taken)
        if taken == 0:
            break
    shape = grid_shape(angles) if sweeps else None
    for _ in range(sweeps if shape else 0):
        top = engine.topn(best_all, topn)
        nb = torch.from_numpy(neighbour_rows(top, *shape).astype(np.int32)).to(dev)
        nat.call("vx_filter_unscored", int(nb.numel()), _p(nb), _p(scored), _p(sel), cnt_p, _st())
        n = int(cnt[0])
        if n == 0:
            break
        score_new(n)
        nscored += n
    top = engine.topn(best_all, topn)
    return top, nscored, Rc


def recall(a, b) -> float:
    a, b = list(a), list(b)
    return len(set(a) & set(b)) / max(1, len(b))
