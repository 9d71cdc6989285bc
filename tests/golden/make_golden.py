"""Writes the committed golden fixtures of this directory.

The reference ships no test vectors and cannot be imported here (its numba,
pyfftw and mrcfile dependencies are absent), so the fixtures are produced by
the oracle (oracle/vesper_ref.py, oracle/fastvec_ref.c), which is itself pinned
by closed forms, brute force and invariants in tests/test_oracle_pins.py.
They freeze small seeded cases of every hot-path function so that the GPU
tests compare against fixed data, and the known-answer tables (FFTW-fast
lengths, Euler row counts) the reference's behaviour fixes.

Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import vesper_ref as ref          # noqa: E402
from oracle.build_c import fastvec_c, load as load_c  # noqa: E402


def maps(rng, n, density):
    data = rng.uniform(0.1, 1.0, (n, n, n)) * (rng.uniform(size=(n, n, n)) < density)
    vec = rng.standard_normal((n, n, n, 3))
    vec /= np.linalg.norm(vec, axis=-1, keepdims=True)
    vec = (vec * (data != 0)[..., None]).astype(np.float32)
    return data, vec


def grid(data, vec):
    st = ref.map_stats(data)
    return ref.MapGrid(data=data, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=vec,
                       ave=st["ave"], std=st["std"], std_norm_ave=st["std_norm_ave"])


def rot_cases():
    for n, seed in ((9, 11), (10, 12)):
        rng = np.random.default_rng(seed)
        data, vec = maps(rng, n, 0.6)
        angles = np.array([[0, 0, 0], [90, 0, 0], [30, 60, 90], [200, 10, 170], [45, 135, 15]], np.float64)
        od, ov = [], []
        for a in angles:
            v, d = ref.rot_mrc(data, vec, a, fma=True)
            od.append(d)
            ov.append(v)
        np.savez_compressed(os.path.join(HERE, f"rot_mrc_n{n}.npz"), data=data, vec=vec, angles=angles,
                            out_data=np.stack(od), out_vec=np.stack(ov))


def search_cases():
    for mode, n, seed in (("VecProduct", 12, 21), ("CC", 11, 22)):
        rng = np.random.default_rng(seed)
        td, tv = maps(rng, n, 0.35)
        sd, sv = maps(rng, n, 0.35)
        t = grid(td, tv)
        spec = ref.target_spectra(ref.target_channels(t, mode))
        angles = np.concatenate([[[0, 0, 0]], rng.uniform(0, 360, (7, 3))])
        best = {"ref32": [], "ref64": []}
        trans = {"ref32": [], "ref64": []}
        for prec in ("ref32", "ref64"):
            for a in angles:
                nv, nd = ref.rot_mrc(sd, sv, a, prec=prec)
                q = ref.query_channels(nd, nv, mode, t.ave)
                b, tr = ref.find_best_trans_list(ref.fft_correlate(spec, q, prec))
                best[prec].append(float(b))
                trans[prec].append(tr)
        np.savez_compressed(os.path.join(HERE, f"search_{mode}_n{n}.npz"), tdata=td, tvec=tv, sdata=sd,
                            svec=sv, ave=t.ave, angles=angles, best32=np.array(best["ref32"]),
                            best64=np.array(best["ref64"]), trans32=np.array(trans["ref32"], np.int32),
                            trans64=np.array(trans["ref64"], np.int32))


def get_score_case():
    n = 10
    rng = np.random.default_rng(31)
    td, tv = maps(rng, n, 0.5)
    sd, sv = maps(rng, n, 0.5)
    t = grid(td, tv)
    out = {}
    trans = np.array([[0, 0, 0], [3, 9, 1], [7, 2, 5]], np.int32)
    for k in ("Nm", "total", "cc_raw", "pcc_raw", "score"):
        out[k] = []
    for tr in trans:
        e = ref.get_score(t, sd, sv, tuple(int(x) for x in tr), "ref64")
        for k in out:
            out[k].append(e[k])
    np.savez_compressed(os.path.join(HERE, "get_score_n10.npz"), tdata=td, tvec=tv, sdata=sd, svec=sv,
                        ave=t.ave, trans=trans, **{k: np.array(v, np.float64) for k, v in out.items()})


def fastvec_case():
    lib = load_c()
    rng = np.random.default_rng(41)
    s = 16
    src = np.zeros((s, s + 2, s + 1), np.float32)
    m = rng.uniform(size=src.shape) < 0.2
    src[m] = rng.uniform(0.2, 1.5, m.sum()).astype(np.float32)
    so = np.array([0.5, -1.0, 2.0], np.float32)
    do = np.array([-1.3, 0.2, 1.9])
    n, w, dreso = 9, 1.5, 3.0
    d, v = fastvec_c(lib, src, 1.0, so, n, w, do, dreso)
    np.savez_compressed(os.path.join(HERE, "fastvec_n9.npz"), src=src, src_width=np.float32(1.0), src_orig=so,
                        n=n, dest_width=w, dest_orig=do, dreso=dreso, data=d, vec=v)


def interp_case():
    from oracle import interp_ref as ir
    rng = np.random.default_rng(51)
    d = np.zeros((9, 8, 7, 4), np.float32)
    sl = (slice(1, 9, 2), slice(0, 8, 2), slice(1, 7, 2))
    d[sl] = (rng.uniform(0.01, 1.0, d[sl].shape) * (rng.uniform(size=d[sl].shape) < 0.8)).astype(np.float32)
    d[1, 0, 1, 0] = 0.5
    out = d.copy()
    parity = ir.interpolate_loops(out)
    np.savez_compressed(os.path.join(HERE, "interp_9x8x7.npz"), grid=d, out=out, parity=np.array(parity))


def tables():
    fast = [ref.next_fast_len(k) for k in range(1, 401)]
    rows = {str(a): int(len(ref.calc_angle_comb(a))) for a in (30.0, 10.0, 5.0)}
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump({"next_fast_len_1_to_400": fast, "angle_rows": rows,
                   "refine_rows_for_angle_0_0_0": int(len(ref.refine_angles([0.0, 0.0, 0.0])))}, f)
    a30 = ref.calc_angle_comb(30.0)
    F = np.stack([ref.flip_matrix_f32(a).reshape(9) for a in a30])
    np.savez_compressed(os.path.join(HERE, "angle_rows_a30.npz"), angles=a30, flip_f32=F)


if __name__ == "__main__":
    rot_cases()
    search_cases()
    get_score_case()
    fastvec_case()
    interp_case()
    tables()
    print("fixtures written to", HERE)
