"""Prob-mode parity: 3 vector + 4 probability channels per rotation, the two
correlation groups, the alpha-mixed z-score (find_best_trans_mixed,
VESPER_1_prob_new.py:1530-1545), and the end-to-end two-pass search
(search_map_fft_prob, :798-1141) against the oracle."""
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
RTOL = 1e-4


def _maps(rng, n, density=0.4):
    data = rng.uniform(0.1, 1.0, (n, n, n)) * (rng.uniform(size=(n, n, n)) < density)
    vec = rng.standard_normal((n, n, n, 3))
    vec /= np.linalg.norm(vec, axis=-1, keepdims=True)
    vec = (vec * (data != 0)[..., None]).astype(np.float32)
    probs = [(rng.uniform(0, 1, (n, n, n)) * (rng.uniform(size=(n, n, n)) < 0.3)).astype(np.float32)
             for _ in range(4)]
    return data, vec, probs


def _near(got_b, got_t, best, trans, arr, label):
    scale = max(abs(best), 1e-30)
    assert abs(got_b - best) <= RTOL * scale, (label, got_b, best)
    if tuple(got_t) != tuple(trans):
        assert abs(arr[tuple(got_t)] - best) <= RTOL * scale, (label, got_t, trans)


@pytest.mark.parametrize("n", [16, 21])
def test_score_rows_prob_groups_and_mix(gpu, n):
    rng = np.random.default_rng(600 + n)
    td, tv, tp = _maps(rng, n)
    sd, sv, sp = _maps(rng, n)
    angles = np.concatenate([[[0, 0, 0]], rng.uniform(0, 360, (5, 3))])
    spec = ref.target_spectra([tv[..., 0], tv[..., 1], tv[..., 2]] + tp)
    alpha, stats = 0.5, (3.0, 0.2, 5.0, 0.1)      # vstd, vave, pstd, pave
    print("[case] prob", n, flush=True)
    plan = engine.SearchPlan(n, nat.VX_MODE_PROB, batch=4)
    plan.set_target(engine.to_dev(td, torch.float64), engine.to_dev(tv, torch.float32),
                    engine.to_dev(np.stack(tp), torch.float32))
    plan.set_search(engine.to_dev(sd, torch.float64), engine.to_dev(sv, torch.float32),
                    engine.to_dev(np.stack(sp), torch.float32))
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    plan.set_prob_stats(alpha, 1.0, 0.0, 1.0, 0.0, enable=False)
    b1, t1 = plan.score_rows(rot)
    plan.set_prob_stats(alpha, *stats, enable=True)
    b2, t2 = plan.score_rows(rot)
    b1, t1, b2, t2 = b1.cpu().numpy(), t1.cpu().numpy(), b2.cpu().numpy(), t2.cpu().numpy()
    assert np.array_equal(b1, b2[:, :2]) and np.array_equal(t1, t2[:, :2])
    for i, a in enumerate(angles):
        nv, nd, q1, q2, q3, q4 = ref.rot_mrc_prob(sd, sv, *sp, a, prec="ref64")
        rv = ref.fft_correlate(spec[:3], [nv[..., 0], nv[..., 1], nv[..., 2]], "ref64")
        rp = ref.fft_correlate(spec[3:], [q1, q2, q3, q4], "ref64")
        vs, vt = ref.find_best_trans_list(rv)
        ps, pt = ref.find_best_trans_list(rp)
        ms, mt = ref.find_best_trans_mixed(rv, rp, alpha, *stats, prec="ref64")
        mixed = (1 - alpha) * (sum(rv) - stats[1]) / stats[0] + alpha * (sum(rp) - stats[3]) / stats[2]
        _near(b2[i, 0], t2[i, 0], vs, vt, sum(rv), ("vec", i))
        _near(b2[i, 1], t2[i, 1], ps, pt, sum(rp), ("prob", i))
        _near(b2[i, 2], t2[i, 2], ms, mt, mixed, ("mixed", i))


@pytest.fixture(scope="module")
def prob_pair(tmp_path_factory):
    d = tmp_path_factory.mktemp("probpair")
    p = synth.make_config_pair("pair100", 3001)
    rng = np.random.default_rng(3001)
    synth.add_prob_channels(rng, p.a)
    synth.add_prob_channels(rng, p.b)
    return synth.write_pair(p, str(d / "p")), p


def _resample_prob(paths, p, s=7.0, g=16.0):
    """main.py:283-361 through the drop-in functions."""
    pa = np.load(paths["A_prob"]).astype(np.float32)
    pb = np.load(paths["B_prob"]).astype(np.float32)
    m1, m2 = V.mrc_obj(paths["A"]), V.mrc_obj(paths["B"])
    p1 = [V.mrc_obj(paths["A"]) for _ in range(4)]
    p2 = [V.mrc_obj(paths["B"]) for _ in range(4)]
    for k in range(4):
        p1[k].data = np.swapaxes(pa[..., k], 0, 2)
        p2[k].data = np.swapaxes(pb[..., k], 0, 2)
    m1, n1 = V.mrc_set_vox_size(m1, p.a.threshold, s)
    m2, n2 = V.mrc_set_vox_size(m2, p.b.threshold, s)
    r1 = [V.mrc_set_vox_size(m, 0.0, s) for m in p1]
    r2 = [V.mrc_set_vox_size(m, 0.0, s) for m in p2]
    lst = [n1, n2] + [r[1] for r in r1] + [r[1] for r in r2]
    dim = max(m.xdim for m in lst)
    for m in lst:
        if m.xdim != dim:
            m.xdim = m.ydim = m.zdim = dim
            m.orig = m.cent - 0.5 * s * m.xdim
        m.data = np.zeros((dim,) * 3, dtype="float32")
        m.vec = np.zeros((dim,) * 3 + (3,), dtype="float32")
    n1 = V.fastVEC(m1, n1, g)
    n2 = V.fastVEC(m2, n2, g)
    t_p = [V.fastVEC(r[0], r[1], g) for r in r1]
    s_p = [V.fastVEC(r[0], r[1], g) for r in r2]
    return n1, n2, t_p, s_p


def _grid(m):
    return ref.MapGrid(data=np.asarray(m.data), xwidth=np.float32(m.xwidth), orig=m.orig, vec=m.vec,
                       cent=m.cent, xdim=m.xdim, ave=m.ave, std=m.std, std_norm_ave=m.std_norm_ave)


def test_prob_search_matches_oracle(gpu, prob_pair):
    paths, p = prob_pair
    with contextlib.redirect_stdout(io.StringIO()):
        n1, n2, t_p, s_p = _resample_prob(paths, p)
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        got = V.search_map_fft_prob(*t_p, n1, n2, *s_p, ang=30.0, alpha=0.5, TopN=5)
    exp = ref.search_map_fft_prob([_grid(m) for m in t_p], _grid(n1), _grid(n2), [_grid(m) for m in s_p],
                                  ang=30.0, alpha=0.5, TopN=5, prec="ref32")
    out = buf.getvalue()
    assert "DotScore Std=" in out and "ProbScore Std=" in out and out.count("Probability Score=") == 5
    assert [tuple(float(v) for v in r["angle"]) for r in got] == \
           [tuple(float(v) for v in r["angle"]) for r in exp["refined"]]
    for g, e in zip(got, exp["refined"]):
        for k in ("vec", "prob", "mixed"):
            assert tuple(g[k + "_trans"]) == tuple(e[k + "_trans"]), k
            assert abs(g[k + "_score"] - e[k + "_score"]) <= RTOL * abs(e[k + "_score"]), k


def test_cli_prob_runs(gpu, prob_pair):
    paths, p = prob_pair
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        res = cli.main(["prob", "-a", paths["A"], "-npa", paths["A_prob"], "-b", paths["B"],
                        "-npb", paths["B_prob"], "-t", str(p.a.threshold), "-T", str(p.b.threshold),
                        "-s", "7", "-A", "30", "-N", "3", "-alpha", "0.5"])
    assert len(res) == 3 and "###Start Refining###" in buf.getvalue()


def test_rot_and_search_fft_prob_dropin(gpu):
    """The drop-in per-row prob function (VESPER_1_prob_new.py:1462-1527),
    without and with the mixing statistics, against the oracle."""
    n = 14
    rng = np.random.default_rng(660)
    td, tv, tp = _maps(rng, n)
    sd, sv, sp = _maps(rng, n)
    # the reference's target_list: conj(rfftn) of the 3 vector and 4
    # probability channels, complex128 (VESPER_1_prob_new.py:846-860)
    spec = ref.target_spectra([tv[..., 0], tv[..., 1], tv[..., 2]] + tp)
    tl = list(spec)
    alpha, stats = 0.3, (2.0, 0.1, 4.0, 0.2)
    for ang in ([0, 0, 0], [40, 110, 70]):
        print("[case] rot_and_search_fft_prob", ang, flush=True)
        g = V.rot_and_search_fft_prob(sd, sv, *sp, ang, tl, alpha)
        assert g[4] is None and g[5] is None
        gm = V.rot_and_search_fft_prob(sd, sv, *sp, ang, tl, alpha, *stats)
        e = ref.rot_and_search_fft_prob(sd, sv, sp, ang, spec, alpha, stats, prec="ref64")
        rv = sum(ref.fft_correlate(spec[:3], [e[6][..., k] for k in range(3)], "ref64"))
        _near(g[0], g[1], e[0], e[1], rv, "vec")
        assert g[:4] == gm[:4]
        assert abs(gm[4] - e[4]) <= RTOL * max(abs(e[4]), 1e-30) and gm[5] == tuple(e[5])
        assert abs(g[2] - e[2]) <= RTOL * max(abs(e[2]), 1e-30)


def test_best_trans_helpers_match_oracle(gpu):
    """find_best_trans_list / find_best_trans_mixed on caller-held maps: the
    float32 sums, z-scores and first argmax bit-exact against the oracle."""
    rng = np.random.default_rng(670)
    n = 12
    vm = [rng.standard_normal((n, n, n)).astype(np.float32) * 50 for _ in range(3)]
    pm = [rng.uniform(0, 30, (n, n, n)).astype(np.float32) for _ in range(4)]
    vm[0][3, 4, 5] = 500.0                               # a clear maximum
    tie = [np.ones((n, n, n), np.float32) for _ in range(2)]
    for maps in (vm, pm, tie):
        print("[case] sum_argmax", len(maps), flush=True)
        assert V.find_best_trans_list(maps) == ref.find_best_trans_list(maps)
    for args in ((0.5, 3.0, 0.2, 5.0, 0.1), (0.0, 1.0, 0.0, 1.0, 0.0), (1.0, 7.3, -1.1, 2.2, 0.4)):
        print("[case] mixed_argmax", args, flush=True)
        g = V.find_best_trans_mixed(vm, pm, *args)
        e = ref.find_best_trans_mixed(vm, pm, *args, prec="ref32")
        assert g[0] == e[0] and tuple(g[1]) == tuple(e[1]), (g, e)


def test_sco_arr_matches_oracle(gpu):
    n = 13
    rng = np.random.default_rng(680)
    td, tv, _ = _maps(rng, n, 0.5)
    sd, sv, _ = _maps(rng, n, 0.5)
    st = ref.map_stats(td)
    t = ref.MapGrid(data=td, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=tv,
                    ave=st["ave"], std=st["std"], std_norm_ave=st["std_norm_ave"])
    for trans in ((0, 0, 0), (2, 11, 5), (12, 1, 7)):
        print("[case] sco_arr", trans, flush=True)
        g = engine.sco_arr(engine.to_dev(tv, torch.float32), engine.to_dev(sv, torch.float32), trans).cpu().numpy()
        e = ref.get_score(t, sd, sv, trans, "ref32")["sco_arr"]
        print("[sco_arr] exact fraction", np.mean(g == e), flush=True)
        assert np.allclose(g, e, rtol=1e-6, atol=1e-7)
