"""GPU parity of each libvesper_hip kernel against the oracle on seeded inputs.

Bars: integer/index work bit-exact (gather, translations, top-N rows); float
scores within 1e-4 relative of the FP32 oracle (ref32) and of the FP64
oracle (ref64).  A translation may differ from the oracle's only at a
near-tie, which is checked by re-scoring both translations in FP64.
"""
import numpy as np
import pytest
import torch

from oracle import vesper_ref as ref
from oracle.build_c import fastvec_c
from vesper_python_amd import _native as nat
from vesper_python_amd import engine
from vesper_python_amd.geometry import rotation_rows

pytestmark = pytest.mark.gpu
RTOL = 1e-4


def _say(*a):
    print("[case]", *a, flush=True)


def _maps(rng, n, density=0.4):
    data = rng.uniform(0.1, 1.0, (n, n, n)) * (rng.uniform(size=(n, n, n)) < density)
    vec = rng.standard_normal((n, n, n, 3))
    vec /= np.linalg.norm(vec, axis=-1, keepdims=True)
    vec = (vec * (data != 0)[..., None]).astype(np.float32)
    return data, vec


def _dev(a, dt):
    return engine.to_dev(a, dt)


# ---------------------------------------------------------------- gather

@pytest.mark.parametrize("n", [8, 9, 12, 15, 24])
@pytest.mark.parametrize("fma", [True, False])
def test_rotate_map_bit_exact(gpu, n, fma):
    rng = np.random.default_rng(100 + n)
    data, vec = _maps(rng, n, 0.6)
    d, v = _dev(data, torch.float64), _dev(vec, torch.float32)
    angles = [[0, 0, 0], [90, 0, 0], [30, 60, 90]] + rng.uniform(0, 360, (4, 3)).tolist()
    for ang in angles:
        _say("rotate_map", n, fma, ang)
        od, ov, _ = engine.rotate_map(rotation_rows(ang)[0], d, v, flags=0 if fma else nat.VX_FLAG_NO_FMA)
        ev, ed = ref.rot_mrc(data, vec, ang, fma=fma)
        assert np.array_equal(od.cpu().numpy(), ed), ang
        assert np.array_equal(ov.cpu().numpy(), ev), ang


@pytest.mark.parametrize("n", [10, 11])
def test_rotate_map_prob_bit_exact(gpu, n):
    rng = np.random.default_rng(200 + n)
    data, vec = _maps(rng, n, 0.5)
    probs = [rng.uniform(0, 1, (n, n, n)).astype(np.float32) * (rng.uniform(size=(n, n, n)) < 0.5)
             for _ in range(4)]
    d, v = _dev(data, torch.float64), _dev(vec, torch.float32)
    p = _dev(np.stack(probs), torch.float32)
    for ang in ([0, 0, 0], [15, 75, 130], [200, 10, 170]):
        _say("rotate_map prob", n, ang)
        od, ov, op = engine.rotate_map(rotation_rows(ang)[0], d, v, p)
        e = ref.rot_mrc_prob(data, vec, *probs, ang)
        assert np.array_equal(od.cpu().numpy(), e[1])
        assert np.array_equal(ov.cpu().numpy(), e[0])
        opn = op.cpu().numpy()
        for k in range(4):
            assert np.array_equal(opn[k], e[2 + k]), k


@pytest.mark.parametrize("n,mode,fma", [(24, "VecProduct", True), (25, "VecProduct", True),
                                        (25, "VecProduct", False), (33, "CC", True), (30, "Overlap", True),
                                        (27, "PCC", False), (39, "Laplacian", True), (21, "prob", True),
                                        (22, "prob", False), (175, "VecProduct", True)])
def test_query_planes_bit_exact(gpu, n, mode, fma):
    """The planes the search itself correlates (its gather kernel, with the
    support box and the compact source) equal the oracle's rot_mrc /
    rot_mrc_prob followed by the mode's query channels, bit for bit: odd
    edges take the centre planes' two candidates per axis, prob mode the
    probability channels' own masks, and the density sits in a sphere so
    that the box prunes."""
    rng = np.random.default_rng(700 + n + 3 * int(fma))
    c = n / 2
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - c
    inner = np.linalg.norm(g, axis=-1) <= 0.3 * n
    sd, sv = _maps(rng, n, 0.5)
    sd[~inner] = 0
    sv[~inner] = 0
    code = nat.VX_MODE_PROB if mode == "prob" else nat.MODE_CODES[mode]
    probs = None
    if mode == "prob":
        probs = [(rng.uniform(0, 1, (n, n, n)) * (rng.uniform(size=(n, n, n)) < 0.5) * inner).astype(np.float32)
                 for _ in range(4)]
    plan = engine.SearchPlan(n, code, batch=1, flags=0 if fma else nat.VX_FLAG_NO_FMA)
    plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32),
                    _dev(np.stack(probs), torch.float32) if probs else None, 0.3)
    angles = [[0, 0, 0], [90, 0, 0], [30, 60, 90], [200, 10, 170]] if n < 100 else [[30, 60, 90]]
    for ang in angles + (rng.uniform(0, 360, (3, 3)).tolist() if n < 100 else []):
        _say("query planes", n, mode, fma, ang)
        got = plan.query_planes(torch.from_numpy(rotation_rows(ang)[0]).cuda()).cpu().numpy()
        if mode == "prob":
            e = ref.rot_mrc_prob(sd, sv, *probs, ang, fma=fma)
            want = [e[0][..., 0], e[0][..., 1], e[0][..., 2]] + list(e[2:])
        else:
            nv, nd = ref.rot_mrc(sd, sv, ang, fma=fma)
            want = ref.query_channels(nd, nv, mode, 0.3)
        assert got.shape[0] == len(want)
        for k, w in enumerate(want):
            assert np.array_equal(got[k], np.asarray(w).astype(np.float32)), (n, mode, fma, ang, k)
    plan.close()


# ---------------------------------------------------------------- FFT search

def _oracle_rows(tdata, tvec, sdata, svec, angles, mode, prec):
    st = ref.map_stats(tdata)
    t = ref.MapGrid(data=tdata, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=tvec,
                    ave=st["ave"], std=st["std"], std_norm_ave=st["std_norm_ave"])
    spec = ref.target_spectra(ref.target_channels(t, mode))
    out = []
    for a in angles:
        nv, nd = ref.rot_mrc(sdata, svec, a, prec=prec)
        q = ref.query_channels(nd, nv, mode, t.ave)
        res = ref.fft_correlate(spec, q, prec)
        s = np.zeros_like(res[0])
        for r in res:
            s = s + r
        best, trans = ref.find_best_trans_list(res)
        out.append((float(best), trans, s))
    return out, t


def _check_rows(best, trans, o32, o64, label):
    for i, ((b32, t32, _), (b64, t64, s64)) in enumerate(zip(o32, o64)):
        g = float(best[i])
        scale = max(abs(b64), 1e-30)
        assert abs(g - b64) <= RTOL * scale, (label, i, g, b64)
        assert abs(b32 - b64) <= RTOL * scale
        tt = tuple(int(v) for v in trans[i])
        if tt != t64:
            # near-tie: the FP64 correlation at our translation is the max
            assert abs(s64[tt] - b64) <= RTOL * scale, (label, i, tt, t64, s64[tt], b64)


@pytest.mark.parametrize("n", [12, 15, 16, 20, 22, 26, 35, 39])
def test_score_rows_vec_mode(gpu, n):
    rng = np.random.default_rng(300 + n)
    td, tv = _maps(rng, n, 0.3)
    sd, sv = _maps(rng, n, 0.3)
    angles = np.concatenate([[[0, 0, 0], [90, 90, 0]], rng.uniform(0, 360, (10, 3))])
    _say("score_rows vec", n)
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=4)
    plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32))
    plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32))
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    best, trans = plan.score_rows(rot)
    best = best[:, 0].cpu().numpy()
    trans = trans[:, 0].cpu().numpy()
    o32, _ = _oracle_rows(td, tv, sd, sv, angles, "VecProduct", "ref32")
    o64, _ = _oracle_rows(td, tv, sd, sv, angles, "VecProduct", "ref64")
    _check_rows(best, trans, o32, o64, f"vec n={n}")


def test_score_rows_self_copy_finds_shift(gpu):
    n = 24
    rng = np.random.default_rng(5)
    td, tv = _maps(rng, n, 0.3)
    shift = (5, 17, 2)
    sd = np.roll(td, shift, axis=(0, 1, 2))
    sv = np.roll(tv, shift, axis=(0, 1, 2))
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=1)
    plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32))
    plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32))
    best, trans = plan.score_rows(torch.from_numpy(rotation_rows([[0, 0, 0]])).cuda())
    assert tuple(trans[0, 0].cpu().tolist()) == shift
    assert abs(best[0, 0].item() / n ** 3 - np.sum(tv.astype(np.float64) ** 2)) < 1e-4 * np.sum(tv ** 2)


@pytest.mark.parametrize("mode,code", [("Overlap", nat.VX_MODE_OVERLAP), ("CC", nat.VX_MODE_CC),
                                       ("PCC", nat.VX_MODE_PCC), ("Laplacian", nat.VX_MODE_LAPLACIAN)])
def test_score_rows_scalar_modes(gpu, mode, code):
    n = 18
    rng = np.random.default_rng(400 + code)
    td, tv = _maps(rng, n, 0.35)
    sd, sv = _maps(rng, n, 0.35)
    angles = rng.uniform(0, 360, (6, 3))
    o32, t = _oracle_rows(td, tv, sd, sv, angles, mode, "ref32")
    o64, _ = _oracle_rows(td, tv, sd, sv, angles, mode, "ref64")
    _say("score_rows", mode)
    plan = engine.SearchPlan(n, code, batch=3)
    plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32), None, t.ave)
    plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32), None, t.ave)
    best, trans = plan.score_rows(torch.from_numpy(rotation_rows(angles)).cuda())
    _check_rows(best[:, 0].cpu().numpy(), trans[:, 0].cpu().numpy(), o32, o64, mode)


# ---------------------------------------------------------------- top-N

def test_topn_stable_order_with_ties(gpu):
    rng = np.random.default_rng(9)
    for nrows, k in ((10, 3), (5000, 10), (24624, 10), (70000, 100), (3000, 1024)):
        s = rng.integers(0, 50, nrows).astype(np.float32)   # many exact ties
        _say("topn", nrows, k)
        rows = engine.topn(torch.from_numpy(s).cuda(), k)
        exp = ref.stable_topn(list(s), k)
        assert rows.tolist() == exp


# ---------------------------------------------------------------- fastVEC + stats

@pytest.mark.parametrize("case", [(20, 10, 2.0, 4.0), (30, 12, 2.5, 6.0), (16, 9, 1.5, 3.0)])
def test_fastvec_matches_c_oracle(gpu, fastvec_ref, case):
    s, n, w, dreso = case
    rng = np.random.default_rng(s)
    src = np.zeros((s, s + 2, s + 1), np.float32)
    m = rng.uniform(size=src.shape) < 0.2
    src[m] = rng.uniform(0.2, 1.5, m.sum()).astype(np.float32)
    so = np.array([0.5, -1.0, 2.0], np.float32)
    do = np.array([-1.3, 0.2, 1.9])
    _say("fastvec", case)
    d_t, v_t = engine.fastvec(src, np.float32(1.0), so, n, w, do, dreso)
    d_c, v_c = fastvec_c(fastvec_ref, src, 1.0, so, n, w, do, dreso)
    d_g = d_t.cpu().numpy()
    v_g = v_t.cpu().numpy()
    assert np.array_equal(d_g != 0, d_c != 0) and (d_c != 0).sum() > 10
    assert np.allclose(d_g, d_c, rtol=1e-12, atol=1e-300)
    assert np.max(np.abs(v_g - v_c)) <= 4e-7
    st = engine.map_stats(d_t)
    e = ref.map_stats(d_c)
    assert st["Nact"] == e["Nact"]
    for k in ("dsum", "ave", "std", "std_norm_ave"):
        assert np.isclose(st[k], e[k], rtol=1e-12), k


# ---------------------------------------------------------------- get_score

@pytest.mark.parametrize("trans", [(0, 0, 0), (3, 9, 1), (11, 2, 7)])
def test_get_score_fp64(gpu, trans):
    n = 14
    rng = np.random.default_rng(sum(trans) + 7)
    td, tv = _maps(rng, n)
    sd, sv = _maps(rng, n)
    st = ref.map_stats(td)
    t = ref.MapGrid(data=td, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=tv,
                    ave=st["ave"], std=st["std"], std_norm_ave=st["std_norm_ave"])
    e = ref.get_score(t, sd, sv, trans, "ref64")
    _say("get_score", trans)
    r = engine.get_score(_dev(td, torch.float64), _dev(tv, torch.float32), st["ave"],
                         _dev(sd, torch.float64), _dev(sv, torch.float32), trans)
    assert r["Nm"] == e["Nm"] and r["total"] == e["total"]
    assert np.isclose(r["cc_raw"], e["cc_raw"], rtol=1e-12)
    assert np.isclose(r["pcc_raw"], e["pcc_raw"], rtol=1e-10, atol=1e-12)
    assert np.isclose(r["score"], e["score"], rtol=1e-12)
    e32 = ref.get_score(t, sd, sv, trans, "ref32")
    assert abs(r["score"] - e32["score"]) <= 1e-5 * max(1.0, abs(e32["score"]))


@pytest.mark.parametrize("n", [24, 25, 40])
def test_support_box_pruning_is_exact(gpu, n):
    """A query whose density sits in a small central ball: the pruned search
    (gather and forward FFT restricted to the support box) returns bit-for-bit
    the unpruned results, and both match the oracle."""
    rng = np.random.default_rng(500 + n)
    td, tv = _maps(rng, n, 0.3)
    sd, sv = _maps(rng, n, 0.6)
    c = n / 2
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - c
    outside = np.linalg.norm(g, axis=-1) > n / 5
    sd[outside] = 0
    sv[outside] = 0
    angles = np.concatenate([[[0, 0, 0]], rng.uniform(0, 360, (7, 3))])
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    res = []
    for flags in (0, nat.VX_FLAG_NO_PRUNE):
        _say("prune", n, flags)
        plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=3, flags=flags)
        plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32))
        plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32))
        info = plan.info()
        if flags == 0:
            assert info["box_width"] < n
        else:
            assert info["box_lo"] == 0 and info["box_width"] == n
        b, t = plan.score_rows(rot)
        res.append((b.cpu().numpy(), t.cpu().numpy()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    o32, _ = _oracle_rows(td, tv, sd, sv, angles, "VecProduct", "ref32")
    o64, _ = _oracle_rows(td, tv, sd, sv, angles, "VecProduct", "ref64")
    _check_rows(res[0][0][:, 0], res[0][1][:, 0], o32, o64, f"prune n={n}")


@pytest.mark.parametrize("n", [24, 27, 40])
def test_pencil_passes_match_slab(gpu, n):
    """The pencil passes (separate z and y transforms, used where a slab
    exceeds LDS) give the slab path's results, and both match the oracle."""
    rng = np.random.default_rng(700 + n)
    td, tv = _maps(rng, n, 0.3)
    sd, sv = _maps(rng, n, 0.3)
    c = n / 2
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - c
    outside = np.linalg.norm(g, axis=-1) > n / 3
    sd[outside] = 0
    sv[outside] = 0
    angles = np.concatenate([[[0, 0, 0]], rng.uniform(0, 360, (6, 3))])
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    res = []
    for pencil in (False, True):
        _say("pencil", n, pencil)
        plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=3)
        plan.tune(pencil=pencil)
        plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32))
        plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32))
        b, t = plan.score_rows(rot)
        res.append((b.cpu().numpy(), t.cpu().numpy()))
    assert np.allclose(res[0][0], res[1][0], rtol=2e-6)
    o32, _ = _oracle_rows(td, tv, sd, sv, angles, "VecProduct", "ref32")
    o64, _ = _oracle_rows(td, tv, sd, sv, angles, "VecProduct", "ref64")
    for b, t in res:
        _check_rows(b[:, 0], t[:, 0], o32, o64, f"pencil n={n}")


def test_large_grid_uses_pencils(gpu):
    """N = 250 (slab of 250 x 126 complex exceeds LDS): scores and shift of a
    translated self-copy through the pencil path."""
    n = 250
    rng = np.random.default_rng(11)
    td = np.zeros((n, n, n))
    tv = np.zeros((n, n, n, 3), np.float32)
    blob = rng.uniform(0.2, 1.0, (40, 40, 40)) * (rng.uniform(size=(40, 40, 40)) < 0.3)
    td[100:140, 100:140, 100:140] = blob
    v = rng.standard_normal((40, 40, 40, 3))
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    tv[100:140, 100:140, 100:140] = (v * (blob != 0)[..., None]).astype(np.float32)
    shift = (7, 240, 19)
    sd = np.roll(td, shift, axis=(0, 1, 2))
    sv = np.roll(tv, shift, axis=(0, 1, 2))
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=1)
    plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32))
    plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32))
    best, trans = plan.score_rows(torch.from_numpy(rotation_rows([[0, 0, 0]])).cuda())
    assert tuple(trans[0, 0].cpu().tolist()) == shift
    exp = np.sum(tv.astype(np.float64) ** 2)
    assert abs(best[0, 0].item() / n ** 3 - exp) < 1e-4 * exp


# ------------------------------------------------ split-bf16 MFMA x pass

@pytest.mark.parametrize("n,flags", [(16, 0), (24, nat.VX_FLAG_NO_PRUNE), (35, 0), (40, 0)])
def test_mfma_xpass_matches_oracle(gpu, n, flags):
    """x-axis DFTs on MFMA with split bf16 operands and fp32 accumulation:
    best scores within 1e-4 of the FP64 oracle; a translation may differ only
    at an FP64 near-tie."""
    rng = np.random.default_rng(900 + n)
    td, tv = _maps(rng, n, 0.3)
    sd, sv = _maps(rng, n, 0.5)
    c = n / 2
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - c
    outside = np.linalg.norm(g, axis=-1) > n / 3
    sd[outside] = 0
    sv[outside] = 0
    angles = np.concatenate([[[0, 0, 0]], rng.uniform(0, 360, (7, 3))])
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    _say("mfma xpass", n, flags)
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=3, flags=flags)
    plan.tune(xpass_variant=2)
    plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32))
    plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32))
    b, t = plan.score_rows(rot)
    o32, _ = _oracle_rows(td, tv, sd, sv, angles, "VecProduct", "ref32")
    o64, _ = _oracle_rows(td, tv, sd, sv, angles, "VecProduct", "ref64")
    _check_rows(b[:, 0].cpu().numpy(), t[:, 0].cpu().numpy(), o32, o64, f"mfma n={n}")


def test_mfma_xpass_cc_mode(gpu):
    n = 22
    rng = np.random.default_rng(950)
    td, tv = _maps(rng, n, 0.4)
    sd, sv = _maps(rng, n, 0.4)
    angles = rng.uniform(0, 360, (6, 3))
    o32, t = _oracle_rows(td, tv, sd, sv, angles, "CC", "ref32")
    o64, _ = _oracle_rows(td, tv, sd, sv, angles, "CC", "ref64")
    _say("mfma xpass cc", n)
    plan = engine.SearchPlan(n, nat.VX_MODE_CC, batch=4)
    plan.tune(xpass_variant=2)
    plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32), None, t.ave)
    plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32), None, t.ave)
    b, tr = plan.score_rows(torch.from_numpy(rotation_rows(angles)).cuda())
    _check_rows(b[:, 0].cpu().numpy(), tr[:, 0].cpu().numpy(), o32, o64, "mfma cc")


def test_mfma_xpass_agrees_with_fft_at_n96(gpu):
    """Grid edge 96 with a pruned support box: the MFMA x pass and the FFT x
    pass give best scores within 1e-4 relative of each other on every row."""
    n = 96
    rng = np.random.default_rng(960)
    td, tv = _maps(rng, n, 0.05)
    sd, sv = _maps(rng, n, 0.2)
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - n / 2
    outside = np.linalg.norm(g, axis=-1) > n / 4
    sd[outside] = 0
    sv[outside] = 0
    rot = torch.from_numpy(rotation_rows(rng.uniform(0, 360, (24, 3)))).cuda()
    res = []
    for variant in (1, 2):
        _say("mfma vs fft n=96 variant", variant)
        plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=8)
        plan.tune(xpass_variant=variant)
        plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32))
        plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32))
        b, t = plan.score_rows(rot)
        res.append((b[:, 0].cpu().numpy().astype(np.float64), t[:, 0].cpu().numpy()))
    (bf, tf), (bm, tm) = res
    rel = np.abs(bm - bf) / np.abs(bf).max()
    print("[mfma] max rel diff", rel.max(), "trans agree", np.mean(np.all(tf == tm, axis=1)), flush=True)
    assert rel.max() <= RTOL


# ------------------------------------------- per-channel correlation helpers

@pytest.mark.parametrize("n", [12, 15, 20])
def test_fft_search_best_dot_matches_oracle(gpu, n):
    """fft_search_best_dot (VESPER_1_prob_new.py:464-491) through
    vx_correlate: every channel's correlation map within 1e-4 of the FP64
    oracle's largest value."""
    from vesper_python_amd import VESPER_1_prob_new as V
    rng = np.random.default_rng(800 + n)
    td, tv = _maps(rng, n, 0.4)
    sd, sv = _maps(rng, n, 0.4)
    spec = ref.target_spectra([tv[..., k] for k in range(3)])
    q = [sv[..., k] for k in range(3)]
    _say("best_dot", n)
    got = V.fft_search_best_dot(spec, q, None, None, None, None, None)
    e64 = ref.fft_correlate(spec, q, "ref64")
    for g, e in zip(got, e64):
        assert g.dtype == np.float32 and g.shape == (n, n, n)
        assert np.max(np.abs(g - e)) <= RTOL * np.max(np.abs(e))
    b, t = V.find_best_trans_list(got)
    eb, et = ref.find_best_trans_list(e64)
    s64 = sum(e64)
    assert abs(b - eb) <= RTOL * abs(eb)
    assert t == et or abs(s64[t] - eb) <= RTOL * abs(eb)


def _target(td, tv):
    st = ref.map_stats(td)
    return ref.MapGrid(data=td, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=tv, ave=st["ave"],
                       std=st["std"], std_norm_ave=st["std_norm_ave"], xdim=td.shape[0])


@pytest.mark.parametrize("mode", ["Overlap", "CC", "PCC", "Laplacian"])
def test_fft_search_score_trans_1d_modes(gpu, mode):
    """fft_search_score_trans_1d (VESPER_1_prob_new.py:529-565): the query
    transform on the device, the correlation with the caller's spectrum and
    its argmax, against the oracle's channels (pinned by closed forms)."""
    from vesper_python_amd import VESPER_1_prob_new as V
    n = 14
    rng = np.random.default_rng(850)
    td, tv = _maps(rng, n, 0.4)
    sd, sv = _maps(rng, n, 0.4)
    t = _target(td, tv)
    spec = ref.target_spectra(ref.target_channels(t, mode))
    qch = ref.query_channels(sd, sv, mode, t.ave)[0]
    _say("score_trans_1d", mode)
    b, tr = V.fft_search_score_trans_1d(spec[0], sd, None, None, None, None, mode, t.ave)
    e = ref.fft_correlate(spec, [np.asarray(qch, np.float32)], "ref64")[0]
    eb = float(e.max())
    assert abs(b - eb) <= RTOL * abs(eb)
    assert abs(e[tr] - eb) <= RTOL * abs(eb)


@pytest.mark.parametrize("mode", ["VecProduct", "Overlap", "CC", "PCC", "Laplacian"])
def test_mode_planes_bit_exact(gpu, mode):
    """vx_mode_planes = the oracle's query channels rounded to float32, bit
    for bit (formed in float64 first, Laplacian in the reference's order)."""
    n = 13
    rng = np.random.default_rng(860)
    sd, sv = _maps(rng, n, 0.5)
    sd[2, 3, 4] = -0.5                                  # a negative voxel stays 0
    ave = 0.41
    code = nat.MODE_CODES[mode]
    got = engine.mode_planes(code, _dev(sd, torch.float64), _dev(sv, torch.float32), None, ave).cpu().numpy()
    want = [np.asarray(c, np.float64).astype(np.float32) for c in ref.query_channels(sd, sv, mode, ave)]
    assert got.shape[0] == len(want)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


@pytest.mark.parametrize("mode", ["VecProduct", "Overlap", "CC", "PCC", "Laplacian"])
def test_rot_and_search_fft_dropin(gpu, mode):
    """rot_and_search_fft (VESPER_1_prob_new.py:1418-1459) with the
    reference's target_list of conjugated complex128 spectra: rotated maps
    bit-exact, the score within 1e-4 of ref32 and ref64 (CC/PCC scaled by the
    target's norms), the translation the oracle's or an FP64 near-tie."""
    from vesper_python_amd import VESPER_1_prob_new as V
    n = 15
    rng = np.random.default_rng(870)
    td, tv = _maps(rng, n, 0.4)
    sd, sv = _maps(rng, n, 0.4)
    t = _target(td, tv)
    tl = ref.target_spectra(ref.target_channels(t, mode))      # built once, as the reference does
    for ang in ([0, 0, 0], [30, 100, 60], [200, 15, 170]):
        _say("rot_and_search_fft", mode, ang)
        gs, gt, gv, gd = V.rot_and_search_fft(sd, sv, ang, tl, t, mode)
        es, et, ev, ed = ref.rot_and_search_fft(sd, sv, ang, tl, t, mode, "ref32")
        e64, et64, _, _ = ref.rot_and_search_fft(sd, sv, ang, tl, t, mode, "ref64")
        assert np.array_equal(gv, ev) and np.array_equal(gd, ed) and gd.dtype == ed.dtype
        assert abs(gs - es) <= RTOL * abs(es) and abs(gs - e64) <= RTOL * abs(e64)
        if tuple(gt) != tuple(et64):
            corr = sum(ref.fft_correlate(tl, ref.query_channels(ed, ev, mode, t.ave), "ref64"))
            assert abs(corr[tuple(gt)] - corr[tuple(et64)]) <= RTOL * abs(corr[tuple(et64)])
        if mode in ("CC", "PCC"):
            assert isinstance(gs, np.float64)
        else:
            assert isinstance(gs, np.float32)


def test_slab_shapes_identical_at_n175(gpu):
    """N = 175 (radices 5 and 7): the 1024-thread slab kernels compiled for
    those radices and the generic 512-thread ones, and the x pass in 256-,
    128- and 64-thread workgroups (tiles of 8, 4 and 2 columns), give
    bit-identical scores, and the rows match the FP64 oracle."""
    n = 175
    rng = np.random.default_rng(1750)
    c = n / 2
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - c
    inner = np.linalg.norm(g, axis=-1) <= 30
    td, tv = _maps(rng, n, 0.02)
    sd, sv = _maps(rng, n, 0.3)
    sd[~inner] = 0
    sv[~inner] = 0
    angles = np.array([[0, 0, 0], [30, 60, 90], [200, 10, 170]], np.float64)
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    res = []
    for nt, tile in ((0, 8), (512, 8), (0, 4), (0, 2)):
        _say("slab shape n=175", nt, "x tile", tile)
        plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=2)
        plan.tune(slab_threads=nt, xpass_variant=3, xpass_tile=tile)
        plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32))
        plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32))
        b, t = plan.score_rows(rot)
        res.append((b.cpu().numpy(), t.cpu().numpy()))
    for b, t in res[1:]:
        assert np.array_equal(res[0][0], b) and np.array_equal(res[0][1], t)
    o32, _ = _oracle_rows(td, tv, sd, sv, angles[:2], "VecProduct", "ref32")
    o64, _ = _oracle_rows(td, tv, sd, sv, angles[:2], "VecProduct", "ref64")
    _check_rows(res[0][0][:2, 0], res[0][1][:2, 0], o32, o64, "n=175")


@pytest.mark.parametrize("n,mode", [(175, nat.VX_MODE_VEC), (175, nat.VX_MODE_PROB),
                                    (128, nat.VX_MODE_VEC), (128, nat.VX_MODE_CC),
                                    (96, nat.VX_MODE_VEC), (104, nat.VX_MODE_VEC), (110, nat.VX_MODE_CC),
                                    (91, nat.VX_MODE_VEC), (162, nat.VX_MODE_PROB),
                                    (24, nat.VX_MODE_CC), (45, nat.VX_MODE_VEC), (64, nat.VX_MODE_PROB),
                                    (80, nat.VX_MODE_VEC), (77, nat.VX_MODE_VEC),
                                    (200, nat.VX_MODE_VEC), (182, nat.VX_MODE_CC), (250, nat.VX_MODE_VEC),
                                    (270, nat.VX_MODE_VEC), (308, nat.VX_MODE_PROB)])
def test_static_plans_identical(gpu, n, mode):
    """Grid edges with compile-time stage plans (175 = 5 5 7, 128 = 8 8 2,
    96 = 8 4 3, 104 = 8 13, 110 = 2 5 11, 91 = 7 13 with a runtime x pass,
    162 = 2 3 3 3 3; in the 8 x 512 slab shape 24, 45, 64, 80, 77 = 7 11;
    in the 24 x 1024 shape 200 = 8 5 5, 182 = 2 7 13; 250 = 2 5 5 5 on the
    pencils with a compile-time x pass in 8-column tiles, 270 = 2 3 3 3 5
    and 308 = 4 7 11 in 4-column tiles):
    the slab and x-pass kernels built on them give the
    runtime-descriptor kernels' scores and translations, in the one-group
    modes and in prob mode (two groups, mixed score).  The FFT code may
    contract a*b + c differently in the two forms, so scores agree to fp32
    rounding (1e-5 relative here); at 175 they are bit-identical.  At 128 the
    rows also match the FP64 oracle."""
    rng = np.random.default_rng(5700 + n + mode)
    c = n / 2
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - c
    inner = np.linalg.norm(g, axis=-1) <= 0.2 * n
    td, tv = _maps(rng, n, 0.03)
    sd, sv = _maps(rng, n, 0.3)
    sd[~inner] = 0
    sv[~inner] = 0
    angles = np.array([[0, 0, 0], [30, 60, 90], [200, 10, 170], [10, 350, 45]], np.float64)
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    prob = None
    if mode == nat.VX_MODE_PROB:
        tp = rng.uniform(0.0, 1.0, (4, n, n, n)).astype(np.float32) * (td != 0)
        sp = rng.uniform(0.0, 1.0, (4, n, n, n)).astype(np.float32) * (sd != 0)
        prob = (_dev(tp, torch.float32), _dev(sp, torch.float32))
    res = []
    for static in (True, False):
        _say("static plans", n, mode, static)
        plan = engine.SearchPlan(n, mode, batch=2)
        plan.tune(static_plans=static, xpass_variant=3, xpass_tile=8 if n <= 256 else 4)
        plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32), prob[0] if prob else None, 0.3)
        plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32), prob[1] if prob else None, 0.3)
        if prob:
            plan.set_prob_stats(0.4, 2.0, 0.1, 3.0, 0.2, enable=True)
        b, t = plan.score_rows(rot)
        res.append((b.cpu().numpy(), t.cpu().numpy()))
    if n == 175:
        assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    else:
        assert np.allclose(res[0][0], res[1][0], rtol=1e-5, atol=0)
        assert np.array_equal(res[0][1], res[1][1])
    if n == 128 and mode == nat.VX_MODE_VEC:
        o32, _ = _oracle_rows(td, tv, sd, sv, angles[:2], "VecProduct", "ref32")
        o64, _ = _oracle_rows(td, tv, sd, sv, angles[:2], "VecProduct", "ref64")
        _check_rows(res[0][0][:2, 0], res[0][1][:2, 0], o32, o64, "n=128")


@pytest.mark.parametrize("n,mode", [(175, nat.VX_MODE_VEC), (175, nat.VX_MODE_PROB), (175, nat.VX_MODE_CC),
                                    (48, nat.VX_MODE_VEC), (64, nat.VX_MODE_PROB), (77, nat.VX_MODE_CC),
                                    (96, nat.VX_MODE_VEC), (128, nat.VX_MODE_PROB), (162, nat.VX_MODE_VEC),
                                    (104, nat.VX_MODE_VEC), (182, nat.VX_MODE_CC), (200, nat.VX_MODE_PROB),
                                    (208, nat.VX_MODE_VEC), (250, nat.VX_MODE_VEC), (270, nat.VX_MODE_CC),
                                    (308, nat.VX_MODE_PROB)])
def test_fused_slabs_identical(gpu, n, mode):
    """The slab kernels that read their first stage from HBM, fold the
    half-spectrum unpack / pack into the next stage and reduce (or store)
    their last stage from registers give the same scores and
    translations to the separate load / unpack / store passes (bit-identical
    at 175; elsewhere the FFT code may contract a*b + c differently in the
    two kernels, so scores agree to fp32 rounding), in every slab
    shape (8 x 512: 48, 64, 77; 16 x 1024: 96..175; 24 x 1024: 182, 200), with
    odd and even edges, and in prob mode, whose mixing pass has the inverse
    slab write both real maps.  Above 200 the same folds run in the pencil
    passes (forward z, y both ways, inverse z with the argmax)."""
    rng = np.random.default_rng(9100 + n + mode)
    c = n / 2
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - c
    inner = np.linalg.norm(g, axis=-1) <= 0.22 * n
    td, tv = _maps(rng, n, 0.03)
    sd, sv = _maps(rng, n, 0.3)
    sd[~inner] = 0
    sv[~inner] = 0
    angles = np.array([[0, 0, 0], [30, 60, 90], [200, 10, 170], [10, 350, 45], [95, 275, 5]], np.float64)
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    prob = None
    if mode == nat.VX_MODE_PROB:
        tp = rng.uniform(0.0, 1.0, (4, n, n, n)).astype(np.float32) * (td != 0)
        sp = rng.uniform(0.0, 1.0, (4, n, n, n)).astype(np.float32) * (sd != 0)
        prob = (_dev(tp, torch.float32), _dev(sp, torch.float32))
    res = []
    for fused in (True, False):
        _say("fused slabs", n, mode, fused)
        plan = engine.SearchPlan(n, mode, batch=2)
        plan.tune(slab_fused=fused)
        plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32), prob[0] if prob else None, 0.3)
        plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32), prob[1] if prob else None, 0.3)
        if prob:
            plan.set_prob_stats(0.4, 2.0, 0.1, 3.0, 0.2, enable=True)
        b, t = plan.score_rows(rot)
        res.append((b.cpu().numpy(), t.cpu().numpy()))
        plan.close()
    if n == 175:
        assert np.array_equal(res[0][0], res[1][0]), (n, mode, np.abs(res[0][0] - res[1][0]).max())
    else:
        assert np.allclose(res[0][0], res[1][0], rtol=1e-6, atol=0), (n, mode)
    assert np.array_equal(res[0][1], res[1][1]), (n, mode)


@pytest.mark.parametrize("n,mode", [(208, nat.VX_MODE_VEC), (308, nat.VX_MODE_PROB), (343, nat.VX_MODE_CC)])
def test_pencil_chunks_identical(gpu, n, mode):
    """The fused inverse z pass walks several chunks of line pairs per
    workgroup and publishes one argmax atomic; every chunk count gives
    bit-identical scores and translations, also where prob mixing reads the
    real maps it writes."""
    rng = np.random.default_rng(9300 + n)
    c = n / 2
    g = np.indices((n, n, n), dtype=np.float32).transpose(1, 2, 3, 0) - c
    inner = np.linalg.norm(g, axis=-1) <= 0.2 * n
    del g
    td, tv = _maps(rng, n, 0.02)
    sd, sv = _maps(rng, n, 0.3)
    sd[~inner] = 0
    sv[~inner] = 0
    angles = np.array([[0, 0, 0], [30, 60, 90], [200, 10, 170]], np.float64)
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    prob = None
    if mode == nat.VX_MODE_PROB:
        tp = rng.uniform(0.0, 1.0, (4, n, n, n)).astype(np.float32) * (td != 0)
        sp = rng.uniform(0.0, 1.0, (4, n, n, n)).astype(np.float32) * (sd != 0)
        prob = (_dev(tp, torch.float32), _dev(sp, torch.float32))
    plan = engine.SearchPlan(n, mode, batch=3)
    plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32), prob[0] if prob else None, 0.3)
    plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32), prob[1] if prob else None, 0.3)
    if prob:
        plan.set_prob_stats(0.4, 2.0, 0.1, 3.0, 0.2, enable=True)
    res = {}
    for chunks in (1, 2, 3, 0, 1000):
        _say("pencil chunks", n, mode, chunks)
        plan.tune(pencil_chunks=chunks)
        b, t = plan.score_rows(rot)
        res[chunks] = (b.cpu().numpy(), t.cpu().numpy())
    plan.close()
    for chunks, (b, t) in res.items():
        assert np.array_equal(b, res[1][0]), (n, mode, chunks)
        assert np.array_equal(t, res[1][1]), (n, mode, chunks)


@pytest.mark.parametrize("n,mode", [(35, nat.VX_MODE_VEC), (24, nat.VX_MODE_CC)])
def test_xpass_variants_identical(gpu, n, mode):
    """The FFT x-pass kernels (register, LDS, register-prefetch; several
    tiles) give bit-identical scores and translations."""
    rng = np.random.default_rng(1100 + n)
    td, tv = _maps(rng, n, 0.3)
    sd, sv = _maps(rng, n, 0.3)
    rot = torch.from_numpy(rotation_rows(rng.uniform(0, 360, (10, 3)))).cuda()
    res = []
    for variant, tile in ((0, 0), (1, 4), (1, 8), (3, 4), (3, 8)):
        _say("x-pass variant", n, variant, tile)
        plan = engine.SearchPlan(n, mode, batch=4)
        plan.tune(xpass_variant=variant, xpass_tile=tile)
        plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32), None, 0.3)
        plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32), None, 0.3)
        b, t = plan.score_rows(rot)
        res.append((b.cpu().numpy(), t.cpu().numpy()))
    for b, t in res[1:]:
        assert np.array_equal(b, res[0][0]) and np.array_equal(t, res[0][1])


def test_xpass_variants_identical_on_pencil_grid(gpu):
    """N = 270, where the slab exceeds LDS and the pencil passes run: the
    default x pass (register-prefetch, 4-column tiles), the LDS x pass and
    2-column register tiles give bit-identical scores and translations."""
    n = 270
    rng = np.random.default_rng(1270)
    td, tv = _maps(rng, n, 0.3)
    sd, sv = _maps(rng, n, 0.3)
    rot = torch.from_numpy(rotation_rows(rng.uniform(0, 360, (4, 3)))).cuda()
    res = []
    for variant, tile in ((None, None), (1, 4), (3, 2)):
        _say("x-pass variant on a pencil grid", n, variant, tile)
        plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=2)
        if variant is not None:
            plan.tune(xpass_variant=variant, xpass_tile=tile)
        plan.set_target(_dev(td, torch.float64), _dev(tv, torch.float32), None, 0.3)
        plan.set_search(_dev(sd, torch.float64), _dev(sv, torch.float32), None, 0.3)
        b, t = plan.score_rows(rot)
        res.append((b.cpu().numpy(), t.cpu().numpy()))
        plan.close()
    for b, t in res[1:]:
        assert np.array_equal(b, res[0][0]) and np.array_equal(t, res[0][1])


def test_correlate_through_pencils_n250(gpu):
    """vx_correlate at N = 250, where the slab exceeds LDS and the pencil
    passes run: one channel's correlation map within 1e-4 of the FP64
    oracle's largest value."""
    from vesper_python_amd import VESPER_1_prob_new as V
    n = 250
    rng = np.random.default_rng(2500)
    t = np.zeros((n, n, n), np.float32)
    q = np.zeros((n, n, n), np.float32)
    t[90:150, 100:160, 80:140] = rng.standard_normal((60, 60, 60)) * (rng.uniform(size=(60, 60, 60)) < 0.2)
    q[100:140, 95:135, 110:150] = rng.standard_normal((40, 40, 40)) * (rng.uniform(size=(40, 40, 40)) < 0.2)
    spec = ref.target_spectra([t])
    _say("correlate pencils", n)
    got = V.fft_search_best_dot(spec, [q], None, None, None, None, None)[0]
    exp = ref.fft_correlate(spec, [q], "ref64")[0]
    assert np.max(np.abs(got - exp)) <= RTOL * np.max(np.abs(exp))
