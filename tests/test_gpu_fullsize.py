"""Parity at the benchmark's full size: the pair200 workload (grid edge 175,
24,624 rows at -A 10), in every orig mode and in prob mode.  The whole
exhaustive search runs on the GPU; the oracle (the reference's float32
arithmetic, ref32) re-scores

* the GPU's top-10 rows and the next ten: same translations, scores within
  1e-4, and rows 11-20 do not beat the top ten;
* an independent strided sample of rows, chosen without looking at the GPU
  scores: each within 1e-4 with the oracle's translation (or an exact tie),
  and no sampled row beats the GPU's 10th score -- a row the GPU under-scored
  out of its top ten would show here;

and the best row against FP64 (ref64).  Prob mode runs four probability inputs per map:
Dirichlet(0.5) channels on the density mask, Gaussian-smoothed and
renormalised, read from .npy and smoothed by fastVEC as the prob command
does (main.py:302-361)."""
import contextlib
import io
import os
import tempfile

import numpy as np
import pytest
import torch

from oracle import vesper_ref as ref
from oracle.build_c import fastvec_c
from oracle_pool import OraclePool
from vesper_python_amd import _native as nat
from vesper_python_amd import engine, synth
from vesper_python_amd import VESPER_1_prob_new as V
from vesper_python_amd import main as cli
from vesper_python_amd.geometry import calc_angle_comb, rotation_rows

pytestmark = pytest.mark.gpu
RTOL = 1e-4
SAMPLE_STRIDE = 1539         # 17 rows of 24,624, spread over every z


@pytest.fixture(scope="module")
def pair200():
    pair = synth.make_config_pair("pair200", 2001)
    tmp = tempfile.mkdtemp(prefix="vx_full_")
    paths = synth.write_pair(pair, os.path.join(tmp, "p"))
    with contextlib.redirect_stdout(io.StringIO()):
        m1, m2, n1, n2 = cli._resample_pair(paths["A"], paths["B"], pair.a.threshold, pair.b.threshold, 2.0, 16.0)
        n1 = V.fastVEC(m1, n1, 16.0)
        n2 = V.fastVEC(m2, n2, 16.0)
    assert int(n1.xdim) == 175
    return pair, paths, n1, n2


def _tgrid(n1):
    return ref.MapGrid(data=np.asarray(n1.data, np.float64), xwidth=np.float32(n1.xwidth), orig=n1.orig,
                       vec=n1.vec, ave=n1.ave, std=n1.std, std_norm_ave=n1.std_norm_ave, xdim=int(n1.xdim))


def _gpu_search(n1, n2, mode, angles):
    code = nat.MODE_CODES[mode]
    plan = engine.SearchPlan(int(n1.xdim), code, batch=32)
    plan.set_target(*n1._dev, None, n1.ave)
    plan.set_search(*n2._dev, None, n1.ave)
    best, trans = plan.score_rows(torch.from_numpy(rotation_rows(angles)).cuda())
    rows = engine.topn(best[:, 0].contiguous(), 20)
    return best[:, 0].cpu().numpy(), trans[:, 0].cpu().numpy(), rows


def _fp64_sum(tgrid, spec, n2, angle, mode):
    nv, nd = ref.rot_mrc(np.asarray(n2.data, np.float64), n2.vec, angle, prec="ref64")
    return sum(ref.fft_correlate(spec, ref.query_channels(nd, nv, mode, tgrid.ave), "ref64"))


def _check_search(n1, n2, mode):
    angles = calc_angle_comb(10.0)
    print("[case] pair200 search", mode, len(angles), flush=True)
    b, t, rows = _gpu_search(n1, n2, mode, angles)
    tgrid = _tgrid(n1)
    spec = ref.target_spectra(ref.target_channels(tgrid, mode))
    sample = list(range(0, len(angles), SAMPLE_STRIDE))
    want = list(dict.fromkeys([int(r) for r in rows] + sample))
    with OraclePool(np.asarray(n2.data, np.float64), n2.vec, spec, tgrid.ave) as pool:
        oracle = pool.v_rows(want, angles, mode, gts={r: t[r] for r in want}, label="pair200")
    ties = 0
    for r in rows[:10]:
        ob, ot, _ = oracle[int(r)]
        assert abs(b[r] - ob) <= RTOL * abs(ob), (mode, r, b[r], ob)
        gt = tuple(int(v) for v in t[r])
        if gt != ot:
            # allowed only at an FP64 near-tie of the two translations
            s64 = _fp64_sum(tgrid, spec, n2, angles[int(r)], mode)
            assert abs(s64[gt] - s64[ot]) <= RTOL * abs(s64[ot]), (mode, r, gt, ot, s64[gt], s64[ot])
            ties += 1
    print("[top10]", mode, "translation near-ties", ties, flush=True)
    s10 = min(oracle[int(r)][0] for r in rows[:10])
    for r in rows[10:]:
        assert oracle[int(r)][0] <= s10 + RTOL * abs(s10), (mode, r, oracle[int(r)][0], s10)
    worst = 0.0
    for r in sample:
        ob, ot, at = oracle[r]
        rel = abs(b[r] - ob) / abs(ob)
        worst = max(worst, rel)
        assert rel <= RTOL, (mode, r, b[r], ob)
        gt = tuple(int(v) for v in t[r])
        assert gt == ot or abs(at - ob) <= RTOL * abs(ob), (mode, r, gt, ot)
        assert ob <= s10 + RTOL * abs(s10) or r in set(int(x) for x in rows[:10]), (mode, r, ob, s10)
    print("[sample]", mode, len(sample), "rows, worst rel", worst, flush=True)
    return b, t, rows, tgrid, spec, angles


def test_pair200_top10_and_sample_against_oracle(gpu, pair200):
    _, _, n1, n2 = pair200
    b, t, rows, tgrid, spec, angles = _check_search(n1, n2, "VecProduct")
    # FP64 error of the best row's score
    r0 = int(rows[0])
    nv, nd = ref.rot_mrc(np.asarray(n2.data, np.float64), n2.vec, angles[r0], prec="ref64")
    b64, t64 = ref.find_best_trans_list(ref.fft_correlate(spec, ref.query_channels(nd, nv, "VecProduct"), "ref64"))
    print("[fp64] best row", r0, "gpu", float(b[r0]), "fp64", float(b64), "rel",
          abs(float(b[r0]) - float(b64)) / abs(float(b64)), flush=True)
    assert abs(b[r0] - b64) <= RTOL * abs(b64) and tuple(int(v) for v in t[r0]) == tuple(t64)


@pytest.mark.parametrize("mode", ["Overlap", "CC", "PCC"])
def test_pair200_scalar_modes_against_oracle(gpu, pair200, mode):
    """Modes O, C and P on the full-size pair."""
    _, _, n1, n2 = pair200
    _check_search(n1, n2, mode)


def test_pair200_prob_mode_dirichlet_generator(gpu, pair200, fastvec_ref):
    """Prob mode: four Dirichlet(0.5) channels per map (smoothed, renormalised,
    zero off the density mask), stored as the (nz, ny, nx, 4) .npy the prob
    command reads, resampled and smoothed by fastVEC (checked against the C
    oracle), then the two-pass mixed search; the top 10 by mixed score and
    a strided sample against the oracle on the same channels."""
    pair, paths, _, _ = pair200
    rng = np.random.default_rng(3001)
    synth.add_prob_channels(rng, pair.a)
    synth.add_prob_channels(rng, pair.b)
    tmp = tempfile.mkdtemp(prefix="vx_prob_")
    pp = synth.write_pair(pair, os.path.join(tmp, "p"))
    print("[case] prob resample + fastVEC of 10 maps", flush=True)
    with contextlib.redirect_stdout(io.StringIO()):
        n1, n2, n1p, n2p = cli.resample_prob(pp["A"], pp["B"], pp["A_prob"], pp["B_prob"], pair.a.threshold,
                                             pair.b.threshold, 2.0, 16.0)
    n = int(n1.xdim)
    assert n == 175
    # channel smoothing = the oracle's fastVEC of the thresholded channel
    ch = np.ascontiguousarray(np.swapaxes(pair.a.prob[..., 2], 0, 2), np.float32)
    ch = np.where(ch < 0, 0, ch).astype(np.float32)
    # (the oracle's exact-order loop takes minutes for all 175 planes: the
    # eight central planes, which cross the whole density, are checked)
    x0, x1 = n // 2 - 4, n // 2 + 4
    d_c, _ = fastvec_c(fastvec_ref, ch, np.float32(pair.a.width), pair.a.origin.astype(np.float32), n, 2.0,
                       n1p[2].orig, 16.0, planes=(x0, x1))
    assert np.count_nonzero(d_c[x0:x1]) > 100
    assert np.allclose(np.asarray(n1p[2].data, np.float64)[x0:x1], d_c[x0:x1].astype(np.float32), rtol=1e-6, atol=0)
    tp = np.stack([np.asarray(m.data, np.float32) for m in n1p])
    sp = np.stack([np.asarray(m.data, np.float32) for m in n2p])
    angles = calc_angle_comb(10.0)
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    plan = engine.SearchPlan(n, nat.VX_MODE_PROB, batch=16)
    plan.set_target(*n1._dev, engine.to_dev(tp, torch.float32), 0.0)
    plan.set_search(*n2._dev, engine.to_dev(sp, torch.float32), 0.0)
    print("[case] prob pass 1", flush=True)
    plan.set_prob_stats(0.5, 1.0, 0.0, 1.0, 0.0, enable=False)
    b1, _ = plan.score_rows(rot)
    b1 = b1.double().cpu().numpy() / n ** 3
    vave, vstd = float(b1[:, 0].mean()), float(b1[:, 0].std())
    pave, pstd = float(b1[:, 1].mean()), float(b1[:, 1].std())
    alpha = 0.5
    print("[case] prob pass 2 (mixed)", flush=True)
    # the reference mixes the unscaled correlation maps with the rd3-scaled
    # statistics (VESPER_1_prob_new.py:920-926, :1530-1545)
    plan.set_prob_stats(alpha, vstd, vave, pstd, pave, enable=True)
    best, trans = plan.score_rows(rot)
    rows = engine.topn(best[:, 2].contiguous(), 10)
    b = best.cpu().numpy()
    t = trans.cpu().numpy()
    tv = np.asarray(n1.vec, np.float32)
    spec = ref.target_spectra([tv[..., 0], tv[..., 1], tv[..., 2]] + [tp[k] for k in range(4)])
    sdata = np.asarray(n2.data, np.float64)
    sample = list(range(0, len(angles), 4 * SAMPLE_STRIDE))
    top = [int(x) for x in rows]
    want = list(dict.fromkeys(top + sample))
    with OraclePool(sdata, n2.vec, spec, prob=[sp[k] for k in range(4)]) as pool:
        oracle = pool.prob_rows(want, angles, alpha, vstd, vave, pstd, pave, label="prob")
    s10 = min(oracle[r][0] for r in top)
    for r in want:
        ms, mt, vs, vt, ps = oracle[r]
        scale = max(abs(float(ms)), 1.0)
        assert abs(b[r, 2] - ms) <= RTOL * scale, (r, b[r, 2], ms)
        assert abs(b[r, 0] - vs) <= RTOL * abs(vs) and abs(b[r, 1] - ps) <= RTOL * abs(ps), (r, b[r], vs, ps)
        if r in top:
            assert tuple(int(v) for v in t[r, 2]) == tuple(mt), (r, t[r, 2], mt)
            assert tuple(int(v) for v in t[r, 0]) == tuple(vt), (r, t[r, 0], vt)
        else:
            assert float(ms) <= s10 + RTOL * max(abs(s10), 1.0), (r, ms, s10)