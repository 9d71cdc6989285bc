"""Prob-mode search on the device (search_map_fft_prob,
VESPER_1_prob_new.py:798-1141): 3 vector + 4 probability channels per
rotation, two groups of correlations, the pass-1 statistics and the
alpha-mixed z-score of find_best_trans_mixed (:1530-1545) fused into the
argmax pass."""
from __future__ import annotations

from datetime import datetime
from pathlib import Path

import numpy as np
import torch

from . import _native as nat
from . import engine
from .geometry import calc_angle_comb, convert_trans, euler2rot, refine_rows, rotation_rows


def _dev_prob(maps):
    """[4][n^3] float32 device planes of four probability maps."""
    planes = []
    for m in maps:
        dv = getattr(m, "_dev", None)
        if dv is not None:
            planes.append(dv[0].float())
        else:
            planes.append(engine.to_dev(np.asarray(m.data, dtype=np.float32), torch.float32))
    return torch.stack(planes).contiguous()


class ProbSearch:
    def __init__(self, tp, target, search, sp, batch=None):
        from .VESPER_1_prob_new import SEARCH_BATCH, _dev_maps
        self.n = int(target.xdim)
        self.tdata, self.tvec = _dev_maps(target)
        self.sdata, self.svec = _dev_maps(search)
        self.tprob = _dev_prob(tp)
        self.sprob = _dev_prob(sp)
        from .batch import auto_batch
        self.plan = engine.SearchPlan(self.n, nat.VX_MODE_PROB, batch or SEARCH_BATCH or auto_batch(self.n, 7))
        self.plan.set_target(self.tdata, self.tvec, self.tprob, 0.0)
        self.plan.set_search(self.sdata, self.svec, self.sprob, 0.0)
        self.rd3 = 1.0 / (self.n ** 3)

    def score(self, angles):
        rot = torch.from_numpy(rotation_rows(angles)).to(self.tdata.device)
        best, trans = self.plan.score_rows(rot)
        return best, trans

    def rescore(self, angles):
        """FP64 (raw scores [R, G], trans [R, G, 3]) of the rows: vector,
        probability and (once mixing is on) mixed groups (vx_rescore_f64)."""
        rot = torch.from_numpy(rotation_rows(angles)).to(self.tdata.device)
        sc, tr, _ = self.plan.rescore_f64(rot, (self.tdata, self.tvec, self.tprob),
                                          (self.sdata, self.svec, self.sprob))
        return sc.cpu().numpy(), tr.cpu().numpy()

    def rotated(self, angle):
        od, ov, op = engine.rotate_map(rotation_rows(angle)[0], self.sdata, self.svec, self.sprob)
        return od, ov, op


def _row(angle, b, t, rd3, mixed):
    r = {"angle": angle,
         "vec_score": np.float64(b[0]) * rd3, "vec_trans": tuple(int(v) for v in t[0]),
         "prob_score": np.float64(b[1]) * rd3, "prob_trans": tuple(int(v) for v in t[1])}
    if mixed:
        r["mixed_score"] = np.float64(b[2]) * rd3
        r["mixed_trans"] = tuple(int(v) for v in t[2])
    return r


def run_prob_search(mrc_P1, mrc_P2, mrc_P3, mrc_P4, mrc_target, mrc_search, mrc_search_p1, mrc_search_p2,
                    mrc_search_p3, mrc_search_p4, ang, alpha, TopN, vave, vstd, pave, pstd, showPDB, folder):
    from .VESPER_1_prob_new import _print_model_line, get_score, show_vec
    s = ProbSearch((mrc_P1, mrc_P2, mrc_P3, mrc_P4), mrc_target, mrc_search,
                   (mrc_search_p1, mrc_search_p2, mrc_search_p3, mrc_search_p4))
    angle_comb = calc_angle_comb(ang)
    print()
    print("###Start Searching###")
    if not (vave >= 0 and vstd >= 0 and pstd >= 0 and pave >= 0):     # :877
        s.plan.set_prob_stats(alpha, 1.0, 0.0, 1.0, 0.0, enable=False)
        best, _ = s.score(angle_comb)
        b = best.cpu().numpy().astype(np.float64)
        vsc = b[:, 0] * s.rd3
        psc = b[:, 1] * s.rd3
        vstd = np.std(vsc)
        vave = np.mean(vsc)
        pstd = np.std(psc)
        pave = np.mean(psc)
    print()
    print("DotScore Std=", vstd, "DotScore Ave=", vave)
    print("ProbScore Std=", pstd, "ProbScore Ave=", pave)
    print()
    s.plan.set_prob_stats(alpha, vstd, vave, pstd, pave, enable=True)
    best, trans = s.score(angle_comb)
    top = engine.topn(best.reshape(-1), TopN, stride=3, col=2)
    bh = best.cpu().numpy()
    th = trans.cpu().numpy()
    from . import VESPER_1_prob_new as V
    report = {"fp64_recheck": bool(V.FP64_RECHECK), "fp32_top": [int(i) for i in top]}
    if V.FP64_RECHECK and len(top):
        # the rows near the cutoff of the mixed-score ranking, decided in FP64
        near = V.recheck_rows(bh[:, 2].astype(np.float64), top)
        b64, t64 = s.rescore(angle_comb[near])
        pick = V._fp64_top(near, b64[:, 2], TopN)
        top = [int(near[k]) for k in pick]
        report.update(recheck_rows=int(len(near)), fp64_top=top,
                      rank_changes=int(sum(a != b for a, b in zip(report["fp32_top"], top))))
        sorted_topN = [_row(angle_comb[near[k]], b64[k], t64[k], s.rd3, True) for k in pick]
    else:
        sorted_topN = [_row(angle_comb[i], bh[i], th[i], s.rd3, True) for i in top]
    # the reference prints the row dicts under numpy 1.21 (requirements.txt),
    # whose scalar reprs carry no type name
    with np.printoptions(legacy="1.25"):
        for idx, x in enumerate(sorted_topN):
            print("M", str(idx + 1), x)

    if ang > 5.0:
        print("\n###Start Refining###")
        ref_rows = np.concatenate([refine_rows(r["angle"]) for r in sorted_topN], axis=0)
        rbest, rtrans = s.score(ref_rows)
        rtop = engine.topn(rbest.reshape(-1), TopN, stride=3, col=2)
        rb = rbest.cpu().numpy()
        rt = rtrans.cpu().numpy()
        if V.FP64_RECHECK:
            # every refinement row is decided in FP64
            rb, rt = s.rescore(ref_rows)
            report["refine_fp32_top"] = [int(i) for i in rtop]
            rtop = V._fp64_top(np.arange(len(ref_rows)), rb[:, 2], TopN)
            report["refine_fp64_top"] = [int(i) for i in rtop]
        refined_list = [_row(tuple(ref_rows[i]), rb[i], rt[i], s.rd3, True) for i in rtop]
    else:
        refined_list = sorted_topN
    for r in refined_list:
        od, ov, _ = s.rotated(r["angle"])
        r["vec"] = ov.cpu().numpy()
        r["data"] = od.cpu().numpy().astype(np.float32)
        r["_dev"] = (od, ov)

    folder_path = None
    if showPDB:
        folder_path = Path.cwd() / (folder if folder is not None else
                                    "VESPER_RUN_" + datetime.now().strftime('%m%d_%H%M'))
        Path.mkdir(folder_path)
        print("###Writing results to PDB files###")
    else:
        print()
    for i, r in enumerate(refined_list):
        new_trans = convert_trans(mrc_target.cent, mrc_search.cent, euler2rot(r["angle"]), r["mixed_trans"],
                                  mrc_search.xwidth, mrc_search.xdim)
        _print_model_line("\n#", i, r["angle"], new_trans)
        print("Translation=", str(r["vec_trans"]), "Probability Score=", str(r["prob_score"]),
              "Probability Translation=", str(r["prob_trans"]))
        sco_arr = get_score(mrc_target, r["data"], r["vec"], r["vec_trans"], _dev=r["_dev"])
        if showPDB:
            show_vec(mrc_target.orig, r["vec"], r["data"], sco_arr, r["mixed_score"], mrc_search.xwidth,
                     r["mixed_trans"], r["angle"], folder_path, i)
    for r in refined_list:
        r.pop("_dev", None)
    run_prob_search.last_report = report
    return refined_list
