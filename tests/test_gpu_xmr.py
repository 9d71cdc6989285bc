"""The radix-split MFMA x pass (x-pass variant 4): N = N1 x N2, the length-N1
DFTs of the x axis as split-f16 matrix products with per-tile power-of-two
scaling, the radix-N2 stage, twiddles, target product and channel sum in
fp32 registers.

* Against the FP64 oracle (ref64) on small grids of every radix N2 it uses
  (2, 5 and 7 here, N1 from 7 to 28), V and CC modes: best scores within
  1e-5 relative, and a translation that differs from the oracle's is an FP64
  near-tie within 1e-5.
* Against the FFT x pass on the same plan at the grid edges the bench and the
  48-pair workload use (96 = 32 x 3, 160 = 32 x 5, 175 = 25 x 7, 192 = 24 x 8,
  200 = 25 x 8), with a narrow support box (one MFMA K-step in the forward
  DFTs) and the full grid (two): best scores within 1e-5 of the largest.
* The FP64 re-scoring takes its candidates from the FFT x pass whichever x
  pass the search runs, so its results do not depend on the variant.
"""
import numpy as np
import pytest
import torch

from vesper_python_amd import _native as nat
from vesper_python_amd import engine
from vesper_python_amd.geometry import rotation_rows

from test_gpu_kernels import _maps, _oracle_rows

pytestmark = pytest.mark.gpu

XMR_RTOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


def _dev(a, dt):
    return engine.to_dev(a, dt)


def _sphere(n, d, v, radius):
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - n / 2
    outside = np.linalg.norm(g, axis=-1) > radius
    d = d.copy()
    v = v.copy()
    d[outside] = 0
    v[outside] = 0
    return d, v


@pytest.mark.parametrize("n,mode", [(24, "VecProduct"), (35, "VecProduct"), (35, "CC"), (49, "VecProduct"),
                                    (40, "CC"), (56, "VecProduct")])
def test_xmr_matches_oracle(gpu, n, mode):
    rng = np.random.default_rng(1500 + n)
    td, tv = _maps(rng, n, 0.3)
    sd, sv = _sphere(n, *_maps(rng, n, 0.5), n / 3)
    angles = np.concatenate([[[0, 0, 0]], rng.uniform(0, 360, (7, 3))])
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    code = nat.VX_MODE_VEC if mode == "VecProduct" else nat.VX_MODE_CC
    o64, t = _oracle_rows(td, tv, sd, sv, angles, mode, "ref64")
    _say("xmr vs oracle", n, mode)
    plan = engine.SearchPlan(n, code, batch=3)
    plan.tune(xpass_variant=4)
    plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32), None, t.ave)
    plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32), None, t.ave)
    b, tr = plan.score_rows(rot)
    b, tr = b[:, 0].cpu().numpy().astype(np.float64), tr[:, 0].cpu().numpy()
    plan.close()
    worst = 0.0
    for i, (b64, t64, s64) in enumerate(o64):
        scale = max(abs(b64), 1e-30)
        worst = max(worst, abs(b[i] - b64) / scale)
        assert abs(b[i] - b64) <= XMR_RTOL * scale, (n, mode, i, b[i], b64)
        tt = tuple(int(v) for v in tr[i])
        if tt != t64:
            assert abs(s64[tt] - b64) <= XMR_RTOL * scale, (n, mode, i, tt, t64)
    print("[xmr] n", n, mode, "max rel vs ref64", worst, flush=True)


@pytest.mark.parametrize("n", [96, 160, 175, 192, 200])
@pytest.mark.parametrize("wide", [False, True])
def test_xmr_matches_fft_xpass(gpu, n, wide):
    rng = np.random.default_rng(1600 + n + wide)
    td, tv = _maps(rng, n, 0.05)
    sd, sv = _sphere(n, *_maps(rng, n, 0.1), (0.5 if wide else 0.2) * n)
    rot = torch.from_numpy(rotation_rows(rng.uniform(0, 360, (16, 3)))).cuda()
    res = []
    for variant in (None, 4):
        _say("xmr vs fft x pass", n, "wide" if wide else "narrow", variant)
        plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=8)
        if variant is not None:
            plan.tune(xpass_variant=variant)
        plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32))
        plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32))
        b, t = plan.score_rows(rot)
        res.append((b[:, 0].cpu().numpy().astype(np.float64), t[:, 0].cpu().numpy(), plan.info()))
        plan.close()
    (bf, tf, info), (bm, tm, _) = res
    rel = np.abs(bm - bf) / np.abs(bf).max()
    same = np.mean(np.all(tf == tm, axis=1))
    print("[xmr] n", n, "box", info["box_lo"], info["box_width"], "max rel", rel.max(), "translations equal", same,
          flush=True)
    assert rel.max() <= XMR_RTOL
    assert same >= 0.75


def test_rescore_does_not_depend_on_the_xpass(gpu):
    n = 35
    rng = np.random.default_rng(1700)
    td, tv = _maps(rng, n, 0.3)
    sd, sv = _sphere(n, *_maps(rng, n, 0.5), n / 3)
    rot = torch.from_numpy(rotation_rows(rng.uniform(0, 360, (6, 3)))).cuda()
    tgt = (_dev(td, torch.float64), _dev(tv, torch.float32))
    src = (_dev(sd, torch.float64), _dev(sv, torch.float32))
    res = []
    for variant in (None, 4):
        _say("rescore with x pass", variant)
        plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=4)
        if variant is not None:
            plan.tune(xpass_variant=variant)
        plan.set_target(*tgt)
        plan.set_search(*src)
        s, t, _ = plan.rescore_f64(rot, tgt, src)
        res.append((s.cpu().numpy(), t.cpu().numpy()))
        plan.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_xmr_rejects_what_it_cannot_run(gpu):
    plan = engine.SearchPlan(99, nat.VX_MODE_VEC, batch=2)   # 99 = 3 x 33: no N2 <= 8 leaves N1 <= 32
    with pytest.raises(engine.VesperError):
        plan.tune(xpass_variant=4)
    plan.close()
    plan = engine.SearchPlan(24, nat.VX_MODE_PROB, batch=2)  # two correlation groups
    with pytest.raises(engine.VesperError):
        plan.tune(xpass_variant=4)
    plan.close()
