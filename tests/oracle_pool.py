"""Oracle rows in parallel for the full-size parity tests.

The oracle restates the reference's float32 arithmetic in NumPy; one row of
the 175^3 benchmark grid takes seconds (the compensated float32 rotation of
every voxel and three to seven 175^3 FFT pairs).  A full-size test re-scores
tens of rows, so they run on a pool of spawned worker processes (NumPy only:
the workers never touch the GPU and never import torch).  The inputs go to
the workers once, as .npy files they memory-map.

Row functions (each returns plain Python values):
  v_row(i, angle, mode, gt):  orig modes -- the unscaled float32 maximum,
      its first-argmax translation, and the summed map's value at the GPU's
      translation gt (None: not asked);
  prob_row(i, angle, alpha, vstd, vave, pstd, pave):  prob mode -- the mixed
      maximum and translation, and the vector and probability maxima and the
      vector translation.
"""
from __future__ import annotations

import multiprocessing as mp
import os
import shutil
import tempfile
from concurrent.futures import ProcessPoolExecutor

import numpy as np

_S: dict = {}
_THREAD_VARS = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")


def _init(d, root):
    import sys
    if root not in sys.path:
        sys.path.insert(0, root)
    for f in os.listdir(d):
        if f.endswith(".npy"):
            _S[f[:-4]] = np.load(os.path.join(d, f), mmap_mode="r")
    _S["_ave"] = float(np.load(os.path.join(d, "ave.npy")))


def _specs(prefix):
    out, i = [], 0
    while f"{prefix}{i}" in _S:
        out.append(np.asarray(_S[f"{prefix}{i}"]))
        i += 1
    return out


def _v_row(task):
    from oracle import vesper_ref as ref
    i, angle, mode, gt = task
    nv, nd = ref.rot_mrc(np.asarray(_S["data"]), np.asarray(_S["vec"]), np.asarray(angle), prec="ref32")
    res = ref.fft_correlate(_specs("spec"), ref.query_channels(nd, nv, mode, _S["_ave"]), "ref32")
    b, t = ref.find_best_trans_list(res)
    at = None if gt is None else float(_sum(res)[tuple(gt)])
    return i, float(b), tuple(int(v) for v in t), at


def _sum(res):
    s = res[0]
    for r in res[1:]:
        s = s + r
    return s


def _prob_row(task):
    from oracle import vesper_ref as ref
    i, angle, alpha, vstd, vave, pstd, pave = task
    sp = [np.asarray(_S[f"prob{k}"]) for k in range(4)]
    nv, nd, q1, q2, q3, q4 = ref.rot_mrc_prob(np.asarray(_S["data"]), np.asarray(_S["vec"]), *sp,
                                              np.asarray(angle), prec="ref32")
    spec = _specs("spec")
    rv = ref.fft_correlate(spec[:3], [nv[..., 0], nv[..., 1], nv[..., 2]], "ref32")
    rp = ref.fft_correlate(spec[3:], [q1, q2, q3, q4], "ref32")
    ms, mt = ref.find_best_trans_mixed(rv, rp, alpha, vstd, vave, pstd, pave, prec="ref32")
    vs, vt = ref.find_best_trans_list(rv)
    ps, _ = ref.find_best_trans_list(rp)
    return (i, float(ms), tuple(int(v) for v in mt), float(vs), tuple(int(v) for v in vt), float(ps))


class OraclePool:
    """Context manager: a spawn pool whose workers hold the search map
    (data, vec), the target spectra and, in prob mode, the four search
    probability channels."""

    def __init__(self, data, vec, spec, ave=0.0, prob=None, workers=None):
        self.dir = tempfile.mkdtemp(prefix="vx_oracle_pool_")
        np.save(os.path.join(self.dir, "data.npy"), np.ascontiguousarray(data, np.float64))
        np.save(os.path.join(self.dir, "vec.npy"), np.ascontiguousarray(vec))
        np.save(os.path.join(self.dir, "ave.npy"), np.float64(ave if ave is not None else 0.0))
        for k, s in enumerate(spec):
            np.save(os.path.join(self.dir, f"spec{k}.npy"), np.ascontiguousarray(s))
        if prob is not None:
            for k, p in enumerate(prob):
                np.save(os.path.join(self.dir, f"prob{k}.npy"), np.ascontiguousarray(p))
        self.workers = workers or max(1, min(8, (os.cpu_count() or 2) - 1))
        self.ex = None
        self._saved = {}

    def __enter__(self):
        for v in _THREAD_VARS:
            self._saved[v] = os.environ.get(v)
            os.environ[v] = "1"
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        self.ex = ProcessPoolExecutor(max_workers=self.workers, mp_context=mp.get_context("spawn"),
                                      initializer=_init, initargs=(self.dir, root))
        return self

    def __exit__(self, *exc):
        try:
            self.ex.shutdown(wait=True, cancel_futures=True)
        finally:
            for v, old in self._saved.items():
                if old is None:
                    os.environ.pop(v, None)
                else:
                    os.environ[v] = old
            shutil.rmtree(self.dir, ignore_errors=True)
        return False

    def v_rows(self, rows, angles, mode, gts=None, label=""):
        """{row: (best, trans, summed map at gts[row] or None)}"""
        tasks = [(int(r), [float(a) for a in angles[int(r)]], mode,
                  None if gts is None else tuple(int(v) for v in gts[int(r)])) for r in rows]
        out = {}
        for i, b, t, at in self.ex.map(_v_row, tasks):
            print("[case] oracle row", label, mode, i, flush=True)
            out[i] = (b, t, at)
        return out

    def prob_rows(self, rows, angles, alpha, vstd, vave, pstd, pave, label=""):
        """{row: (mixed best, mixed trans, vector best, vector trans, prob best)}"""
        tasks = [(int(r), [float(a) for a in angles[int(r)]], alpha, vstd, vave, pstd, pave) for r in rows]
        out = {}
        for res in self.ex.map(_prob_row, tasks):
            print("[case] oracle prob row", label, res[0], flush=True)
            out[res[0]] = res[1:]
        return out
