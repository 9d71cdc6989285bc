"""Many map pairs and repeated searches: hipGraph-captured search steps.

A search step (every rotation row through gather, FFT correlation and
argmax, then the top-k selection) is captured once per plan into a HIP graph
and replayed for each pair of the same grid edge: the plan's buffers, the
rotation rows and the outputs keep their addresses, so a new pair only
rewrites the plan's target spectra and query source before the replay.  The
support box is pinned to the largest one of the group so that every pair has
the same launch shapes (the union of the group's boxes).  Pairs of uneven size are grouped by grid edge; each
group gets a batch size that keeps one launch's work near that of 32 rotations
at N = 175.  A group of one pair runs eagerly.
"""
from __future__ import annotations

from collections import defaultdict

import numpy as np
import torch

from . import _native as nat
from . import engine
from .geometry import calc_angle_comb, rotation_rows


def auto_batch(n: int, mode_channels: int = 3) -> int:
    """Rotations per launch: 32 at grid edge 175 (measured optimum on
    MI355X), scaled by the grid volume so the workspace stays near 5 GB;
    between 8 and 128."""
    return int(max(8, min(128, round(32 * (175.0 / n) ** 3 * 3.0 / max(1, mode_channels)))))


class GraphedSearch:
    """Captured search step of `plan` over the rows `rot` (device [R, 9])."""

    def __init__(self, plan: engine.SearchPlan, rot: torch.Tensor, k: int = 10):
        self.plan = plan
        self.rot = rot
        R = int(rot.shape[0])
        self.k = min(int(k), R)
        self.best = torch.empty((R, plan.groups), dtype=torch.float32, device=rot.device)
        self.trans = torch.empty((R, plan.groups, 3), dtype=torch.int32, device=rot.device)
        self.rows = torch.empty(max(self.k, 1), dtype=torch.int32, device=rot.device)
        # the top-k key buffers belong to this object, so a replay never
        # touches memory that another call could free or reuse
        self.ws = engine.topn_workspace(R, self.k, rot.device)
        self._step()                      # warm-up outside capture: plan state
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._step()
        torch.cuda.synchronize()

    def _step(self):
        self.plan.score_rows(self.rot, self.best, self.trans)
        G = self.plan.groups
        engine.topn_into(self.best, int(self.rot.shape[0]), self.k, self.rows, self.ws, stride=G, col=G - 1)

    def replay(self):
        self.graph.replay()
        return self.best, self.trans, self.rows


def _mode_scale(target, mode):
    if mode == "CC":
        return 1.0 / target.std ** 2
    if mode == "PCC":
        return 1.0 / target.std_norm_ave ** 2
    return 1.0


def search_pairs(pairs, ang=10.0, mode="VecProduct", topn=10, batch=None, use_graph=True, progress=None):
    """Exhaustive search of every (target, search) pair of resampled maps
    (fastVEC outputs).  Returns per pair a dict(n, scores (1/N^3-scaled,
    float64 [R]), trans [R, 3], top (row ids), angles).  `progress`, if
    given, is called with (pair index, grid edge) after each pair."""
    angles = calc_angle_comb(ang)
    rot_host = rotation_rows(angles)
    dev = engine.require_gpu()
    rot = torch.from_numpy(rot_host).to(dev)
    code = nat.MODE_CODES[mode]
    groups = defaultdict(list)
    for i, (t, s) in enumerate(pairs):
        groups[int(t.xdim)].append(i)
    out = [None] * len(pairs)
    from .VESPER_1_prob_new import _dev_maps
    for n, idxs in sorted(groups.items()):
        plan = engine.SearchPlan(n, code, batch or auto_batch(n))
        dm = {i: (_dev_maps(pairs[i][0]), _dev_maps(pairs[i][1])) for i in idxs}
        lo, hi = n, -1
        for i in idxs:
            plan.set_search(*dm[i][1], None, pairs[i][0].ave or 0.0)
            info = plan.info()
            lo = min(lo, info["box_lo"])
            hi = max(hi, info["box_lo"] + info["box_width"] - 1)
        plan.pin_box(lo, hi - lo + 1)
        gs = None
        # a group of one pair runs eagerly: capturing costs one extra step
        # (the warm-up before capture) and there would be nothing to replay
        graph = use_graph and len(idxs) > 1
        for i in idxs:
            t, s = pairs[i]
            ave = t.ave if t.ave is not None else 0.0
            plan.set_target(*dm[i][0], None, ave)
            plan.set_search(*dm[i][1], None, ave)
            if graph:
                if gs is None:
                    # the warm-up step before capture scores this pair (capture
                    # records without running), so its results are taken as is
                    gs = GraphedSearch(plan, rot, topn)
                    best, trans, rows = gs.best, gs.trans, gs.rows
                else:
                    best, trans, rows = gs.replay()
                top = rows[:gs.k].cpu().numpy()
            else:
                best, trans = plan.score_rows(rot)
                top = engine.topn(best[:, 0].contiguous(), topn)
            b = best[:, 0].cpu().numpy().astype(np.float64) * _mode_scale(t, mode) / n ** 3
            out[i] = dict(n=n, scores=b, trans=trans[:, 0].cpu().numpy(), top=np.asarray(top), angles=angles)
            if progress is not None:
                progress(i, n)
        plan.close()
    return out
