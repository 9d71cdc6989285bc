"""FP64 re-scoring (vx_rescore_f64) and the FP64 re-check of the ranking.

* Per row, the FP64 score and translation equal the FP64 oracle's (ref64:
  FP64 rotation of the vectors, complex128 FFTs, float64 sums) in every mode,
  within 1e-9 relative, translations identical; prob mode's vector,
  probability and mixed groups likewise.
* search_map_fft with the re-check (the default) equals the oracle's search
  with the same re-check (rows near the top-N cutoff and every refinement
  row scored in ref64) on the 100^3/60^3 pair; without it, the float32
  oracle (ref32)."""
import contextlib
import io

import numpy as np
import pytest
import torch

from oracle import vesper_ref as ref
from vesper_python_amd import _native as nat
from vesper_python_amd import engine, synth
from vesper_python_amd import VESPER_1_prob_new as V
from vesper_python_amd import main as cli
from vesper_python_amd.geometry import rotation_rows

pytestmark = pytest.mark.gpu


def _maps(rng, n, density=0.35):
    data = rng.uniform(0.1, 1.0, (n, n, n)) * (rng.uniform(size=(n, n, n)) < density)
    vec = rng.standard_normal((n, n, n, 3))
    vec /= np.linalg.norm(vec, axis=-1, keepdims=True)
    vec = (vec * (data != 0)[..., None]).astype(np.float32)
    return data, vec


def _dev(a, dt):
    return engine.to_dev(a, dt)


@pytest.mark.parametrize("mode", ["VecProduct", "Overlap", "CC", "PCC", "Laplacian"])
@pytest.mark.parametrize("n", [18, 21])
def test_rescore_f64_against_ref64(gpu, mode, n):
    rng = np.random.default_rng(900 + n)
    td, tv = _maps(rng, n)
    sd, sv = _maps(rng, n)
    st = ref.map_stats(td)
    t = ref.MapGrid(data=td, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=tv, ave=st["ave"],
                    std=st["std"], std_norm_ave=st["std_norm_ave"], xdim=n)
    angles = np.concatenate([[[0, 0, 0], [90, 0, 0]], rng.uniform(0, 360, (8, 3))])
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    plan = engine.SearchPlan(n, nat.MODE_CODES[mode], batch=4)
    tdd, tvd, sdd, svd = _dev(td, torch.float64), _dev(tv, torch.float32), _dev(sd, torch.float64), _dev(sv, torch.float32)
    plan.set_target(tdd, tvd, None, t.ave)
    plan.set_search(sdd, svd, None, t.ave)
    print("[case] rescore", mode, n, flush=True)
    sc, tr, nc = plan.rescore_f64(rot, (tdd, tvd), (sdd, svd), t.ave)
    sc, tr, nc = sc[:, 0].cpu().numpy(), tr[:, 0].cpu().numpy(), nc[:, 0].cpu().numpy()
    b32, t32 = plan.score_rows(rot)
    spec = ref.target_spectra(ref.target_channels(t, mode))
    for i, a in enumerate(angles):
        nv, nd = ref.rot_mrc(sd, sv, a, prec="ref64")
        res = ref.fft_correlate(spec, ref.query_channels(nd, nv, mode, t.ave), "ref64")
        b64, t64 = ref.find_best_trans_list(res)
        assert nc[i] >= 1
        assert abs(sc[i] - b64) <= 1e-9 * abs(b64), (mode, i, sc[i], b64)
        gt = tuple(int(v) for v in tr[i])
        if gt != tuple(t64):
            # an exact tie (Overlap counts) that the FFT's rounding broke
            s64 = sum(res)
            assert abs(s64[gt] - b64) <= 1e-9 * abs(b64) and gt < tuple(t64), (mode, i, gt, t64)
        # and the float32 search agrees with it to 1e-4
        assert abs(float(b32[i, 0]) - b64) <= 1e-4 * abs(b64)


def test_rescore_f64_prob_groups(gpu):
    n = 16
    rng = np.random.default_rng(950)
    td, tv = _maps(rng, n)
    sd, sv = _maps(rng, n)
    tp = np.stack([(rng.uniform(0, 1, (n, n, n)) * (rng.uniform(size=(n, n, n)) < 0.3)) for _ in range(4)]).astype(np.float32)
    sp = np.stack([(rng.uniform(0, 1, (n, n, n)) * (rng.uniform(size=(n, n, n)) < 0.3)) for _ in range(4)]).astype(np.float32)
    angles = np.concatenate([[[0, 0, 0]], rng.uniform(0, 360, (6, 3))])
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    plan = engine.SearchPlan(n, nat.VX_MODE_PROB, batch=4)
    tdd, tvd, tpd = _dev(td, torch.float64), _dev(tv, torch.float32), _dev(tp, torch.float32)
    sdd, svd, spd = _dev(sd, torch.float64), _dev(sv, torch.float32), _dev(sp, torch.float32)
    plan.set_target(tdd, tvd, tpd, 0.0)
    plan.set_search(sdd, svd, spd, 0.0)
    alpha, vstd, vave, pstd, pave = 0.5, 3.0e-3, 1.0e-3, 2.0e-3, 4.0e-4
    plan.set_prob_stats(alpha, vstd, vave, pstd, pave, enable=True)
    sc, tr, _ = plan.rescore_f64(rot, (tdd, tvd, tpd), (sdd, svd, spd))
    sc, tr = sc.cpu().numpy(), tr.cpu().numpy()
    spec = ref.target_spectra([tv[..., 0], tv[..., 1], tv[..., 2]] + [tp[k] for k in range(4)])
    for i, a in enumerate(angles):
        print("[case] prob rescore row", i, flush=True)
        vs, vt, ps, pt, ms, mt, _, _ = ref.rot_and_search_fft_prob(sd, sv, tuple(sp), a, spec, alpha,
                                                                    (vstd, vave, pstd, pave), prec="ref64")
        for g, (want, wt) in enumerate(((vs, vt), (ps, pt), (ms, mt))):
            assert abs(sc[i, g] - want) <= 1e-9 * abs(want), (i, g, sc[i, g], want)
            assert tuple(int(v) for v in tr[i, g]) == tuple(wt), (i, g, tr[i, g], wt)


def _grid(m):
    return ref.MapGrid(data=np.asarray(m.data, np.float64), xwidth=np.float32(m.xwidth), orig=m.orig,
                       vec=m.vec, cent=m.cent, xdim=m.xdim, ave=m.ave, std=m.std, std_norm_ave=m.std_norm_ave)


@pytest.mark.parametrize("recheck", [True, False])
@pytest.mark.parametrize("mode", ["VecProduct", "CC"])
def test_search_with_and_without_fp64_recheck(gpu, tmp_path, monkeypatch, recheck, mode):
    p = synth.make_config_pair("pair100", 1001)
    paths = synth.write_pair(p, str(tmp_path / "p"))
    with contextlib.redirect_stdout(io.StringIO()):
        m1, m2, n1, n2 = cli._resample_pair(paths["A"], paths["B"], p.a.threshold, p.b.threshold, 7.0, 16.0)
        n1 = V.fastVEC(m1, n1, 16.0)
        n2 = V.fastVEC(m2, n2, 16.0)
        monkeypatch.setattr(V, "FP64_RECHECK", recheck)
        got = V.search_map_fft(n1, n2, TopN=10, ang=30.0, mode=mode)
    rep = V.search_map_fft.last_report
    exp = ref.search_map_fft(_grid(n1), _grid(n2), TopN=10, ang=30.0, mode=mode, prec="ref32", fp64_recheck=recheck)
    print("[recheck]", mode, recheck, {k: v for k, v in rep.items() if not k.endswith("_top")}, flush=True)
    assert rep["fp64_recheck"] == recheck
    tol = 1e-9 if recheck else 1e-4
    assert [tuple(float(v) for v in r["angle"]) for r in got] == \
           [tuple(float(v) for v in r["angle"]) for r in exp["refined"]]
    for g, e in zip(got, exp["refined"]):
        assert tuple(g["vec_trans"]) == tuple(e["vec_trans"])
        assert abs(g["vec_score"] - e["vec_score"]) <= tol * abs(e["vec_score"])
    if recheck:
        assert rep["recheck_rows"] >= 10 and rep["fp64_top"] == [int(i) for i in exp["topn"]]
