"""Drop-in replacement for the reference module ``VESPER_1_prob_new`` on MI355X.

Same public names, signatures, returned objects and printed lines as
VESPER_1_prob_new.py; every voxel-sized computation (thresholding, fastVEC,
map statistics, rotate-and-gather, FFT correlation, argmax, top-N,
get_score) runs in libvesper_hip's HIP kernels.  pyFFTW plan arguments
(``a, b, c, fft_object, ifft_object``) are accepted and ignored.

The search functions batch every rotation row of a pair through one device
plan instead of rotating and transforming one row at a time.
"""
from __future__ import annotations

import copy
import ctypes
from datetime import datetime
from pathlib import Path

import numpy as np
import torch

from . import _native as nat
from . import engine
from .geometry import (calc_angle_comb as _calc_angle_comb, convert_trans, euler2rot,
                       next_fast_len, refine_rows, rotation_rows)
from .mrc import read_mrc

__all__ = ["mrc_obj", "mrc_set_vox_size", "fastVEC", "doFastVEC", "rot_mrc", "rot_mrc_prob",
           "find_best_trans_list", "find_best_trans_mixed", "laplacian_filter", "search_map_fft",
           "search_map_fft_prob", "rot_and_search_fft", "rot_and_search_fft_prob", "get_score",
           "show_vec", "calc_angle_comb", "euler2rot", "convert_trans"]

# rotations per launch of the search kernels (None: batch.auto_batch of the
# grid edge; overridable by callers)
SEARCH_BATCH = None

# FP64 re-check of the ranking: the rows near the top-N cutoff (within
# FP64_RECHECK_REL of the N-th float32 score, relative to the larger of the
# first and N-th) and every refinement row are re-scored in FP64 on the
# device (vx_rescore_f64), and the top N are taken by those scores.  False:
# rank by the float32 scores alone, as the reference does
# (VESPER_1_prob_new.py:664, :739).
FP64_RECHECK = True
FP64_RECHECK_REL = 1e-4


class mrc_obj:
    """Density data and statistics of an MRC file (VESPER_1_prob_new.py:18-56)."""

    def __init__(self, path):
        m = read_mrc(path)
        data = m.data
        self.xdim = int(m.nx)
        self.ydim = int(m.ny)
        self.zdim = int(m.nz)
        self.xwidth = m.voxel_size[0]
        self.ywidth = m.voxel_size[1]
        self.zwidth = m.voxel_size[2]
        self.cent = np.array([self.xdim * 0.5, self.ydim * 0.5, self.zdim * 0.5])
        self.orig = np.array([m.origin[0], m.origin[1], m.origin[2]])
        self.data = np.swapaxes(copy.deepcopy(data), 0, 2)
        self.dens = data.flatten()
        self.vec = np.zeros((self.xdim, self.ydim, self.zdim, 3), dtype="float32")
        self.dsum = None
        self.Nact = None
        self.ave = None
        self.std_norm_ave = None
        self.std = None


def _threshold_dmax(mrc, th):
    """data < th -> 0 and dmax on the device (VESPER_1_prob_new.py:78-85)."""
    d = engine.to_dev(np.ascontiguousarray(mrc.data, dtype=np.float32), torch.float32)
    out = np.zeros(2, dtype=np.float64)
    sx, sy, sz = (int(s) for s in d.shape)
    nat.call("vx_threshold_dmax", engine._p(d), sx, sy, sz, ctypes.c_float(np.float32(th)),
             out.ctypes.data_as(ctypes.c_void_p), engine._stream())
    return d, float(out[0])


def mrc_set_vox_size(mrc, th=0.00, voxel_size=7.0):
    """Threshold, grid edge and new origin (VESPER_1_prob_new.py:59-118)."""
    if th < 0:
        mrc.dens = mrc.dens - th
        th = 0.0
    mrc.dens[mrc.dens < th] = 0.0
    d, dmax = _threshold_dmax(mrc, th)
    mrc.data = d.cpu().numpy()
    mrc._dev_data = d
    dmax = np.float64(dmax)
    print()
    # the reference prints the voxel-unit extent divided by the voxel width
    print("#dmax=" + str(dmax / mrc.xwidth))
    dmax = dmax * mrc.xwidth
    new_cent = mrc.cent * mrc.xwidth + mrc.orig
    tmp_size = 2 * dmax / voxel_size
    new_xdim = next_fast_len(int(tmp_size))
    new_orig = new_cent - 0.5 * new_xdim * voxel_size
    mrc_new = copy.copy(mrc)
    mrc_new.data = mrc.data.copy()
    mrc_new.vec = mrc.vec.copy()
    mrc_new._dev_data = None
    mrc_new.orig = new_orig
    mrc_new.xdim = new_xdim
    mrc_new.ydim = new_xdim
    mrc_new.zdim = new_xdim
    mrc_new.cent = new_cent
    mrc_new.xwidth = mrc_new.ywidth = mrc_new.zwidth = voxel_size
    print("Nvox= " + str(mrc_new.xdim) + ", " + str(mrc_new.ydim) + ", " + str(mrc_new.zdim))
    print("cent= " + str(new_cent[0]) + ", " + str(new_cent[1]) + ", " + str(new_cent[2]))
    print("ori= " + str(new_orig[0]) + ", " + str(new_orig[1]) + ", " + str(new_orig[2]))
    return mrc, mrc_new


def doFastVEC(src_xwidth, src_orig, src_dims, src_data, dest_xwidth, dest_orig, dest_dims, dest_data,
              dest_vec, dreso=16.0):
    """VESPER_1_prob_new.py:180-257 on the GPU; fills dest_data/dest_vec in place."""
    n = int(dest_dims[0])
    if not (int(dest_dims[1]) == n and int(dest_dims[2]) == n):
        raise ValueError("destination grid must be cubic")
    data_t, vec_t = engine.fastvec(src_data, src_xwidth, src_orig, n, dest_xwidth, dest_orig, dreso)
    dest_data[...] = data_t.cpu().numpy().astype(dest_data.dtype, copy=False)
    dest_vec[...] = vec_t.cpu().numpy()
    return dest_data, dest_vec


def fastVEC(src, dest, dreso=16.0):
    """VESPER_1_prob_new.py:145-177: vector field + statistics of ``dest``.
    Device copies stay attached to ``dest`` for the search."""
    n = int(dest.xdim)
    src_data = getattr(src, "_dev_data", None)
    if src_data is None:
        src_data = src.data
    data_t, vec_t = engine.fastvec(src_data, src.xwidth, src.orig, n, dest.xwidth, dest.orig, dreso)
    out_dtype = np.asarray(dest.data).dtype
    if out_dtype == np.float32:
        # prob mode keeps float32 grids (main.py:345)
        data_t = data_t.float().double()
    st = engine.map_stats(data_t)
    print("#MAP SUM={sum} COUNT={cnt} AVE={ave} STD={std} STD_norm={std_norm}".format(
        sum=st["dsum"], cnt=st["Nact"], ave=st["ave"], std=st["std"], std_norm=st["std_norm_ave"]))
    dest.data = data_t.cpu().numpy().astype(out_dtype, copy=False)
    dest.vec = vec_t.cpu().numpy()
    dest.dsum = st["dsum"]
    dest.Nact = st["Nact"]
    dest.ave = st["ave"]
    dest.std = st["std"]
    dest.std_norm_ave = st["std_norm_ave"]
    dest._dev = (data_t, vec_t)
    return dest


def _dev_maps(m):
    """Device (data float64, vec float32) of a resampled map."""
    dv = getattr(m, "_dev", None)
    if dv is not None and tuple(dv[0].shape) == tuple(np.shape(m.data)):
        return dv
    d = engine.to_dev(np.asarray(m.data, dtype=np.float64), torch.float64)
    v = engine.to_dev(np.asarray(m.vec, dtype=np.float32), torch.float32)
    return d, v


def rot_mrc(orig_mrc_data, orig_mrc_vec, angle):
    """VESPER_1_prob_new.py:279-354 on the GPU: returns (new_vec, new_data)."""
    d = engine.to_dev(np.asarray(orig_mrc_data, dtype=np.float64), torch.float64)
    v = engine.to_dev(np.asarray(orig_mrc_vec, dtype=np.float32), torch.float32)
    od, ov, _ = engine.rotate_map(rotation_rows(angle)[0], d, v)
    return ov.cpu().numpy(), od.cpu().numpy().astype(np.asarray(orig_mrc_data).dtype, copy=False)


def rot_mrc_prob(orig_mrc_data, orig_mrc_vec, mrc_search_p1_data, mrc_search_p2_data, mrc_search_p3_data,
                 mrc_search_p4_data, angle):
    """VESPER_1_prob_new.py:1211-1299 on the GPU: returns 6 arrays."""
    d = engine.to_dev(np.asarray(orig_mrc_data, dtype=np.float64), torch.float64)
    v = engine.to_dev(np.asarray(orig_mrc_vec, dtype=np.float32), torch.float32)
    p = engine.to_dev(np.stack([np.asarray(x, dtype=np.float32) for x in
                                (mrc_search_p1_data, mrc_search_p2_data, mrc_search_p3_data,
                                 mrc_search_p4_data)]), torch.float32)
    od, ov, op = engine.rotate_map(rotation_rows(angle)[0], d, v, p)
    opn = op.cpu().numpy()
    return (ov.cpu().numpy(), od.cpu().numpy().astype(np.asarray(orig_mrc_data).dtype, copy=False),
            opn[0], opn[1], opn[2], opn[3])


def laplacian_filter(arr):
    """VESPER_1_prob_new.py:494-526 on the GPU."""
    a = np.asarray(arr)
    d = engine.to_dev(a.astype(np.float64), torch.float64)
    return engine.laplacian(d).cpu().numpy().astype(a.dtype, copy=False)


def calc_angle_comb(ang_interval):
    """VESPER_1_prob_new.py:1396-1415."""
    return _calc_angle_comb(ang_interval)


_MODE_OF = dict(nat.MODE_CODES)


class _Search:
    """One target/query pair on the device with a plan for a mode."""

    def __init__(self, mrc_target, mrc_search, mode, batch=None):
        if mode not in _MODE_OF:
            raise ValueError(f"unknown mode {mode!r}")
        self.n = int(mrc_target.xdim)
        self.mode = mode
        self.target = mrc_target
        self.search = mrc_search
        self.tdata, self.tvec = _dev_maps(mrc_target)
        self.sdata, self.svec = _dev_maps(mrc_search)
        from .batch import auto_batch
        self.plan = engine.SearchPlan(self.n, _MODE_OF[mode], batch or SEARCH_BATCH or auto_batch(self.n))
        ave = mrc_target.ave if mrc_target.ave is not None else 0.0
        self.plan.set_target(self.tdata, self.tvec, None, ave)
        self.plan.set_search(self.sdata, self.svec, None, ave)
        self.rd3 = 1.0 / (mrc_target.xdim ** 3)
        self.scale = 1.0
        if mode == "CC":
            self.scale = 1.0 / mrc_target.std ** 2
        elif mode == "PCC":
            self.scale = 1.0 / mrc_target.std_norm_ave ** 2

    def score(self, angles):
        """(scores: float64 1/N^3-scaled, trans int [R,3], best device tensor)."""
        rot = torch.from_numpy(rotation_rows(angles)).to(self.tdata.device)
        best, trans = self.plan.score_rows(rot)
        b = best[:, 0].cpu().numpy().astype(np.float64)
        if self.scale != 1.0:
            b = b * self.scale
        return b * self.rd3, trans[:, 0, :].cpu().numpy(), best[:, 0].contiguous()

    def rescore(self, angles):
        """FP64 (scores 1/N^3-scaled, trans [R, 3]) of the rows (vx_rescore_f64)."""
        rot = torch.from_numpy(rotation_rows(angles)).to(self.tdata.device)
        ave = self.target.ave if self.target.ave is not None else 0.0
        sc, tr, _ = self.plan.rescore_f64(rot, (self.tdata, self.tvec), (self.sdata, self.svec), ave)
        b = sc[:, 0].cpu().numpy()
        if self.scale != 1.0:
            b = b * self.scale
        return b * self.rd3, tr[:, 0, :].cpu().numpy()

    def rotated(self, angle):
        od, ov, _ = engine.rotate_map(rotation_rows(angle)[0], self.sdata, self.svec)
        return od, ov


def recheck_rows(scores, top, rel=None):
    """Rows within `rel` (default FP64_RECHECK_REL) of the N-th score of the
    float32 top N (relative to the larger of the first and N-th), in row
    order: the rows whose rank the FP64 re-check re-decides."""
    rel = FP64_RECHECK_REL if rel is None else rel
    s1, sn = scores[top[0]], scores[top[-1]]
    thr = sn - rel * max(abs(sn), abs(s1))
    return np.flatnonzero(np.asarray(scores) >= thr)


def _fp64_top(rows, s64, topn):
    """Top N of `rows` by FP64 score, ties in row order."""
    order = sorted(range(len(rows)), key=lambda k: (-s64[k], int(rows[k])))
    return order[:topn]


def _trans_tuple(t):
    return tuple(int(v) for v in t)


def _print_model_line(prefix, i, angle, new_trans):
    print(prefix + str(i), "Rotation=", "(" + str(angle[0]), str(angle[1]), str(angle[2]) + ")",
          "Translation=", "(" + "{:.3f}".format(new_trans[0]), "{:.3f}".format(new_trans[1]),
          "{:.3f}".format(new_trans[2]) + ")")


def search_map_fft(mrc_target, mrc_search, TopN=10, ang=30, mode="VecProduct", is_eval_mode=False, save_path=".",
                   showPDB=False, folder=None):
    """VESPER_1_prob_new.py:568-795: exhaustive rotational search, top-N,
    +-5 degree refinement (if ang > 5), final scoring.  Returns the refined
    list of dicts (angle, vec_score, vec_trans, vec, data)."""
    if is_eval_mode:
        print("#For Evaluation Mode")
        print("#Please use the same coordinate system and map size for map1 and map2.")
        print("#Example:")
        print("#In Chimera command line: open map1 and map2 as #0 and #1, then type")
        print("#> open map1.mrc")
        print("#> open map2.mrc")
        print("#> vop #1 resample onGrid #0")
        print("#> volume #2 save new.mrc")
        print("#Chimera will generate the resampled map2.mrc as new.mrc")
        _ = get_score(mrc_target, mrc_search.data, mrc_search.vec, [0, 0, 0])
        return None

    s = _Search(mrc_target, mrc_search, mode)
    angle_comb = calc_angle_comb(ang)
    print("###Start Searching###")
    scores, trans, best_dev = s.score(angle_comb)
    ave = np.mean(scores)
    std = np.std(scores)
    print("\nStd= " + str(std) + " Ave= " + str(ave) + "\n")
    top = engine.topn(best_dev, TopN)
    report = {"fp64_recheck": bool(FP64_RECHECK), "fp32_top": [int(i) for i in top]}
    if FP64_RECHECK and len(top):
        near = recheck_rows(scores, top)
        s64, t64 = s.rescore(angle_comb[near])
        pick = _fp64_top(near, s64, TopN)
        top = [int(near[k]) for k in pick]
        report.update(recheck_rows=int(len(near)), fp64_top=top,
                      rank_changes=int(sum(a != b for a, b in zip(report["fp32_top"], top))),
                      trans_changes=int(sum(tuple(t64[k]) != tuple(trans[near[k]]) for k in pick)),
                      max_rel_fp32_fp64=float(np.max(np.abs(s64 - scores[near]) / np.abs(s64))))
        sorted_topN = [{"angle": angle_comb[near[k]], "vec_score": s64[k], "vec_trans": _trans_tuple(t64[k])}
                       for k in pick]
    else:
        sorted_topN = [{"angle": angle_comb[i], "vec_score": scores[i], "vec_trans": _trans_tuple(trans[i])}
                       for i in top]
    for i, r in enumerate(sorted_topN):
        new_trans = convert_trans(mrc_target.cent, mrc_search.cent, euler2rot(r["angle"]), r["vec_trans"],
                                  mrc_search.xwidth, mrc_search.xdim)
        _print_model_line("M", i, r["angle"], new_trans)
    print()

    if ang > 5.0:
        ref_rows = np.concatenate([refine_rows(r["angle"]) for r in sorted_topN], axis=0)
        rs, rt, rbest = s.score(ref_rows)
        rtop = engine.topn(rbest, TopN)
        if FP64_RECHECK:
            # every refinement row is decided in FP64
            rs, rt = s.rescore(ref_rows)
            report["refine_fp32_top"] = [int(i) for i in rtop]
            rtop = _fp64_top(np.arange(len(ref_rows)), rs, TopN)
            report["refine_fp64_top"] = [int(i) for i in rtop]
            report["refine_rank_changes"] = int(sum(a != b for a, b in zip(report["refine_fp32_top"], rtop)))
        refined_list = []
        for i in rtop:
            od, ov = s.rotated(ref_rows[i])
            refined_list.append({"angle": tuple(ref_rows[i]), "vec_score": rs[i], "vec_trans": _trans_tuple(rt[i]),
                                 "vec": ov.cpu().numpy(),
                                 "data": od.cpu().numpy().astype(np.asarray(mrc_search.data).dtype, copy=False),
                                 "_dev": (od, ov)})
    else:
        refined_list = sorted_topN
        for r in refined_list:
            od, ov = s.rotated(r["angle"])
            r["vec"] = ov.cpu().numpy()
            r["data"] = od.cpu().numpy().astype(np.asarray(mrc_search.data).dtype, copy=False)
            r["_dev"] = (od, ov)

    folder_path = None
    if showPDB:
        folder_path = Path.cwd() / (folder if folder is not None else
                                    "VESPER_RUN_" + datetime.now().strftime('%m%d_%H%M'))
        Path.mkdir(folder_path)
        print("\n###Writing results to PDB files###")
    else:
        print()

    for i, r in enumerate(refined_list):
        new_trans = convert_trans(mrc_target.cent, mrc_search.cent, euler2rot(r["angle"]), r["vec_trans"],
                                  mrc_search.xwidth, mrc_search.xdim)
        _print_model_line("\n#", i, r["angle"], new_trans)
        sco_arr = get_score(mrc_target, r["data"], r["vec"], r["vec_trans"], _dev=r.get("_dev"))
        if showPDB:
            show_vec(mrc_target.orig, r["vec"], r["data"], sco_arr, r["vec_score"], mrc_search.xwidth,
                     r["vec_trans"], r["angle"], folder_path, i)
    for r in refined_list:
        r.pop("_dev", None)
    search_map_fft.last_report = report
    return refined_list


def get_score(target_map, search_map_data, search_map_vec, trans, _dev=None):
    """VESPER_1_prob_new.py:1302-1393 in FP64 on the GPU.  Prints the
    Overlap/CC/PCC line and the DOT Score; returns sco_arr (per-voxel dot
    products at the search positions)."""
    tdata, tvec = _dev_maps(target_map)
    if _dev is not None:
        sdata, svec = _dev
    else:
        sdata = engine.to_dev(np.asarray(search_map_data, dtype=np.float64), torch.float64)
        svec = engine.to_dev(np.asarray(search_map_vec, dtype=np.float32), torch.float32)
    r = engine.get_score(tdata, tvec, target_map.ave, sdata, svec, trans)
    std1, pstd1 = target_map.std, target_map.std_norm_ave
    print("Overlap=" + str(float(r["Nm"]) / float(r["total"])) + " " + str(r["Nm"]) + "/" + str(r["total"])
          + " CC=" + str(r["cc_raw"] / (std1 * r["std2"])) + " PCC=" + str(r["pcc_raw"] / (pstd1 * r["pstd2"])))
    print("Score=", np.float64(r["score"]))
    get_score.last = r
    return _sco_arr(tvec, svec, trans, np.asarray(search_map_data).dtype)


def _sco_arr(tvec, svec, trans, dtype):
    """Per-voxel dot products placed at the search positions (sco_arr of
    get_score, VESPER_1_prob_new.py:1362-1370), for the PDB B-factors."""
    out = engine.sco_arr(tvec, svec, trans, torch.float64 if np.dtype(dtype) == np.float64 else torch.float32)
    return out.cpu().numpy()


def show_vec(origin, sampled_mrc_vec, sampled_mrc_data, score_arr, score, sample_width, trans, angle,
             folder_path, rank):
    """CA/CB pseudo-atom PDB of a placed model (VESPER_1_prob_new.py:1144-1208)."""
    dim = sampled_mrc_data.shape[0]
    filename = "R_{:02d}-S_{:>7.3f}.pdb".format(rank, score).replace(" ", "_")
    filepath = Path(folder_path) / filename
    origin = np.array([origin[0], origin[1], origin[2]])
    trans = np.array(trans)
    for a in range(3):
        if trans[a] > 0.5 * dim:
            trans[a] -= dim
    add = origin - trans * sample_width
    idx = np.argwhere(np.asarray(sampled_mrc_data) != 0.0)
    lines = []
    natm, nres = 1, 1
    for x, y, z in idx:
        tmp = np.array([x, y, z]) * sample_width + add
        b = score_arr[x][y][z]
        lines.append("ATOM{:>7d}  CA  ALA{:>6d}    ".format(natm, nres)
                     + "{:8.3f}{:8.3f}{:8.3f}{:6.2f}{:6.2f}".format(tmp[0], tmp[1], tmp[2], 1.0, b))
        natm += 1
        tmp = (np.array([x, y, z]) + sampled_mrc_vec[x][y][z]) * sample_width + add
        lines.append("ATOM{:>7d}  CB  ALA{:>6d}    ".format(natm, nres)
                     + "{:8.3f}{:8.3f}{:8.3f}{:6.2f}{:6.2f}".format(tmp[0], tmp[1], tmp[2], 1.0, b))
        natm += 1
        nres += 1
    with open(filepath, "w") as f:
        f.write("".join(s + "\n" for s in lines))


_CORR_PLANS = {}


def _corr_plan(n):
    """A small plan per grid edge for the per-channel correlation helpers."""
    p = _CORR_PLANS.get(n)
    if p is None:
        p = engine.SearchPlan(n, nat.VX_MODE_CC, batch=1)
        _CORR_PLANS[n] = p
    return p


_SPECTRA = {}


def _spectra_dev(spectra):
    """Device complex64 copy [C, n, n, n/2+1] of a list of conjugated target
    spectra.  The reference builds that list once per search and hands the
    same arrays to every row (VESPER_1_prob_new.py:621-628, :644-655), so the
    upload is kept while the caller passes the same array objects with the
    same contents (identity plus a strided sample of every array)."""
    arrs = [np.asarray(t) for t in spectra]
    key = tuple((id(a), a.shape, a.dtype.str) for a in arrs)
    sample = [a.reshape(-1)[::max(1, a.size // 4096)].copy() for a in arrs]
    hit = _SPECTRA.get(key)
    if hit is not None and all(np.array_equal(x, y) for x, y in zip(hit[1], sample)):
        return hit[2]
    dev = torch.from_numpy(np.stack([a.astype(np.complex64) for a in arrs])).to(engine.require_gpu())
    _SPECTRA.clear()
    _SPECTRA[key] = (arrs, sample, dev)      # holding the arrays keeps their ids unique
    return dev


def _best_trans(maps, n):
    best, i = engine.sum_argmax(list(maps))
    return best, tuple(int(v) for v in np.unravel_index(i, (n, n, n)))


def rot_and_search_fft(data, vec, angle, target_list, mrc_target, mode="VecProduct"):
    """VESPER_1_prob_new.py:1418-1459 for one row on the device: rotate the
    query (vx_rotate_map), form its FFT inputs for the mode (vx_mode_planes),
    correlate them with the caller's conjugated target spectra
    ``target_list`` (``[:3]`` in VecProduct mode, ``[0]`` otherwise;
    vx_correlate), and take the channel sum's first argmax (vx_sum_argmax).
    CC and PCC scores are scaled by 1 / std^2 and 1 / std_norm_ave^2 of
    ``mrc_target`` (:1452-1457).  Returns (score before 1/N^3, trans, rotated
    vec, rotated data)."""
    if mode not in _MODE_OF:
        raise ValueError(f"unknown mode {mode!r}")
    d = engine.to_dev(np.asarray(data, dtype=np.float64), torch.float64)
    v = engine.to_dev(np.asarray(vec, dtype=np.float32), torch.float32)
    n = int(d.shape[0])
    od, ov, _ = engine.rotate_map(rotation_rows(angle)[0], d, v)
    code = _MODE_OF[mode]
    ave = mrc_target.ave if mrc_target.ave is not None else 0.0
    q = engine.mode_planes(code, od, ov, None, ave)
    spec = _spectra_dev(target_list[:3] if mode == "VecProduct" else [target_list[0]])
    corr = _corr_plan(n).correlate(spec, q)
    vec_score, vec_trans = _best_trans(corr, n)
    if mode == "CC":
        vec_score = np.float64(vec_score) * (1.0 / mrc_target.std ** 2)
    if mode == "PCC":
        vec_score = np.float64(vec_score) * (1.0 / mrc_target.std_norm_ave ** 2)
    return (vec_score, vec_trans, ov.cpu().numpy(),
            od.cpu().numpy().astype(np.asarray(data).dtype, copy=False))


def fft_search_best_dot(target_list, query_list, a=None, b=None, c=None, fft_object=None, ifft_object=None):
    """VESPER_1_prob_new.py:464-491 on the device (vx_correlate): the
    unnormalised correlation map irfftn(target * rfftn(query)) per channel.
    ``target_list`` holds conjugated target spectra ([n, n, n/2+1] complex);
    the pyFFTW buffers and plans ``a, b, c, fft_object, ifft_object`` are
    accepted and ignored."""
    q = np.stack([np.asarray(x, dtype=np.float32) for x in query_list])
    n = q.shape[1]
    ts = torch.from_numpy(np.stack([np.asarray(t).astype(np.complex64) for t in target_list]))
    out = _corr_plan(n).correlate(ts.to(engine.require_gpu()), engine.to_dev(q, torch.float32))
    host = out.cpu().numpy()
    return [host[i] for i in range(host.shape[0])]


_MODE_1D = {"Overlap": nat.VX_MODE_OVERLAP, "CC": nat.VX_MODE_CC, "PCC": nat.VX_MODE_PCC,
            "Laplacian": nat.VX_MODE_LAPLACIAN}


def fft_search_score_trans_1d(target_X, search_data, a=None, b=None, fft_object=None, ifft_object=None,
                              mode="CC", ave=None):
    """VESPER_1_prob_new.py:529-565 on the device: the mode's query transform
    (vx_mode_planes), the correlation with ``target_X`` (vx_correlate) and
    its best translation (vx_sum_argmax)."""
    if mode not in _MODE_1D:
        raise ValueError(f"unknown mode {mode!r}")
    d = engine.to_dev(np.asarray(search_data, dtype=np.float64), torch.float64)
    n = int(d.shape[0])
    q = engine.mode_planes(_MODE_1D[mode], d, None, None, 0.0 if ave is None else float(ave))
    ts = torch.from_numpy(np.asarray(target_X).astype(np.complex64)[None])
    out = _corr_plan(n).correlate(ts.to(q.device), q)
    return _best_trans([out[0]], n)


def find_best_trans_list(input_list):
    """Channel sum, max and first argmax (VESPER_1_prob_new.py:357-374),
    reduced on the device (vx_sum_argmax)."""
    arrs = [engine.to_dev(np.asarray(a, dtype=np.float32), torch.float32) for a in input_list]
    best, i = engine.sum_argmax(arrs)
    return best, tuple(int(v) for v in np.unravel_index(i, tuple(arrs[0].shape)))


def find_best_trans_mixed(vec_fft_results, prob_fft_results, alpha, vstd, vave, pstd, pave):
    """VESPER_1_prob_new.py:1530-1545 in float32 on the device (vx_mixed_argmax)."""
    v = [engine.to_dev(np.asarray(a, dtype=np.float32), torch.float32) for a in vec_fft_results]
    p = [engine.to_dev(np.asarray(a, dtype=np.float32), torch.float32) for a in prob_fft_results]
    best, i = engine.mixed_argmax(v, p, alpha, vstd, vave, pstd, pave)
    return best, tuple(int(x) for x in np.unravel_index(i, tuple(v[0].shape)))


def search_map_fft_prob(mrc_P1, mrc_P2, mrc_P3, mrc_P4, mrc_target, mrc_search, mrc_search_p1, mrc_search_p2,
                        mrc_search_p3, mrc_search_p4, ang, alpha=0.0, TopN=10, num_proc=4, vave=-10, vstd=-10,
                        pave=-10, pstd=-10, showPDB=False, folder=None):
    """VESPER_1_prob_new.py:798-1141: vector + probability search with the
    alpha-mixed z-score (two passes unless all four statistics are given)."""
    from .prob import run_prob_search
    return run_prob_search(mrc_P1, mrc_P2, mrc_P3, mrc_P4, mrc_target, mrc_search, mrc_search_p1, mrc_search_p2,
                           mrc_search_p3, mrc_search_p4, ang, alpha, TopN, vave, vstd, pave, pstd, showPDB, folder)


def rot_and_search_fft_prob(data, vec, dp1, dp2, dp3, dp4, angle, target_list, alpha, vstd=None, vave=None,
                            pstd=None, pave=None):
    """VESPER_1_prob_new.py:1462-1527 for one row on the device: rotate the
    query and its four probability maps, each with its own mask
    (vx_rotate_map), correlate the 3 vector components with
    ``target_list[:3]`` and the 4 probability planes with ``target_list[3:]``
    (the caller's conjugated spectra; vx_correlate), and take each group's
    first argmax (vx_sum_argmax); with all four statistics also the
    alpha-mixed z-score's (vx_mixed_argmax, :1530-1545).  Returns
    (vec_score, vec_trans, prob_score, prob_trans, mixed_score, mixed_trans),
    the maxima unscaled; the mixed pair is None without the statistics."""
    if len(target_list) < 7:
        raise ValueError("target_list must hold the 3 vector and 4 probability spectra of the target")
    d = engine.to_dev(np.asarray(data, dtype=np.float64), torch.float64)
    v = engine.to_dev(np.asarray(vec, dtype=np.float32), torch.float32)
    p = engine.to_dev(np.stack([np.asarray(x, dtype=np.float32) for x in (dp1, dp2, dp3, dp4)]), torch.float32)
    n = int(d.shape[0])
    _, ov, op = engine.rotate_map(rotation_rows(angle)[0], d, v, p)
    q = engine.mode_planes(nat.VX_MODE_PROB, None, ov, op)
    corr = _corr_plan(n).correlate(_spectra_dev(target_list[:7]), q)
    vec_score, vec_trans = _best_trans(corr[:3], n)
    prob_score, prob_trans = _best_trans(corr[3:], n)
    mixed_score, mixed_trans = None, None
    if vstd is not None and vave is not None and pstd is not None and pave is not None:
        best, i = engine.mixed_argmax(list(corr[:3]), list(corr[3:]), alpha, vstd, vave, pstd, pave)
        mixed_score, mixed_trans = best, tuple(int(x) for x in np.unravel_index(i, (n, n, n)))
    return vec_score, vec_trans, prob_score, prob_trans, mixed_score, mixed_trans
