"""Captured search steps (HIP graphs) and the multi-pair search: a replayed
step returns the eager step's results bit for bit, replay after replay with
eager work in between, and pairs of uneven grid size searched through
per-edge graphs match the per-pair eager search and the oracle."""
import contextlib
import io

import numpy as np
import pytest
import torch

from oracle import vesper_ref as ref
from vesper_python_amd import _native as nat
from vesper_python_amd import batch, engine, synth
from vesper_python_amd import VESPER_1_prob_new as V
from vesper_python_amd import main as cli
from vesper_python_amd.geometry import rotation_rows

pytestmark = pytest.mark.gpu


def _maps(rng, n, density=0.3):
    data = rng.uniform(0.1, 1.0, (n, n, n)) * (rng.uniform(size=(n, n, n)) < density)
    vec = rng.standard_normal((n, n, n, 3))
    vec /= np.linalg.norm(vec, axis=-1, keepdims=True)
    return data, (vec * (data != 0)[..., None]).astype(np.float32)


def test_graph_replay_matches_eager(gpu):
    n = 24
    rng = np.random.default_rng(21)
    td, tv = _maps(rng, n)
    sd, sv = _maps(rng, n)
    angles = rng.uniform(0, 360, (40, 3))
    rot = torch.from_numpy(rotation_rows(angles)).cuda()
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=8)
    plan.set_target(engine.to_dev(td, torch.float64), engine.to_dev(tv, torch.float32))
    plan.set_search(engine.to_dev(sd, torch.float64), engine.to_dev(sv, torch.float32))
    eb, et = plan.score_rows(rot)
    eb, et = eb.cpu().numpy(), et.cpu().numpy()
    erows = engine.topn(torch.from_numpy(eb[:, 0].copy()).cuda(), 10)
    gs = batch.GraphedSearch(plan, rot, 10)
    for rep in range(4):
        print("[case] replay", rep, flush=True)
        b, t, rows = gs.replay()
        torch.cuda.synchronize()
        assert np.array_equal(b.cpu().numpy(), eb) and np.array_equal(t.cpu().numpy(), et)
        assert np.array_equal(rows.cpu().numpy(), erows)
        plan.score_rows(rot[:8])                # eager work between replays
        torch.cuda.synchronize()


def test_graph_replay_survives_library_scratch_growth(gpu):
    """A captured step's top-k buffers are its own: a larger vx_topn after the
    capture (which grows, frees and reallocates the library's shared scratch)
    leaves later replays bit-identical to the eager step."""
    n = 20
    rng = np.random.default_rng(23)
    td, tv = _maps(rng, n)
    sd, sv = _maps(rng, n)
    rot = torch.from_numpy(rotation_rows(rng.uniform(0, 360, (24, 3)))).cuda()
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=8)
    plan.set_target(engine.to_dev(td, torch.float64), engine.to_dev(tv, torch.float32))
    plan.set_search(engine.to_dev(sd, torch.float64), engine.to_dev(sv, torch.float32))
    eb, _ = plan.score_rows(rot)
    erows = engine.topn(eb[:, 0].contiguous(), 10)
    gs = batch.GraphedSearch(plan, rot, 10)
    for rep, big in enumerate((300_000, 900_000)):
        # a larger selection than any before: the shared scratch is regrown
        scores = torch.rand(big, device="cuda")
        rows = engine.topn(scores, 1024)
        assert np.array_equal(rows[:5], torch.sort(scores, descending=True).indices[:5].cpu().numpy())
        torch.cuda.synchronize()
        print("[case] replay after scratch growth", rep, flush=True)
        _, _, grows = gs.replay()
        torch.cuda.synchronize()
        assert np.array_equal(grows.cpu().numpy(), erows)


def test_topn_ws_equals_topn(gpu):
    rng = np.random.default_rng(24)
    for nrows, k, stride, col in ((5000, 10, 1, 0), (24624, 10, 3, 2), (70000, 1024, 2, 1), (7, 10, 1, 0)):
        s = rng.standard_normal(nrows * stride).astype(np.float32)
        s[rng.integers(0, nrows * stride, nrows // 4)] = 0.5          # ties keep row order
        sd = torch.from_numpy(s).cuda()
        want = engine.topn(sd, k, stride, col)
        ws = engine.topn_workspace(nrows, k)
        rows = torch.empty(max(k, 1), dtype=torch.int32, device="cuda")
        engine.topn_into(sd, nrows, k, rows, ws, stride, col)
        got = rows[:min(k, nrows)].cpu().numpy()
        assert np.array_equal(got, want)
        assert list(got) == ref.stable_topn(list(s[col::stride]), k)
    with pytest.raises(nat.VesperError):
        engine.topn_into(sd, 7, 10, rows, ws[:1], 1, 0)


def _resampled_pair(tmp, seed, a_edge, natoms, s):
    p = synth.make_pair(seed, a_edge=a_edge, natoms=natoms, b_edge=40, nsub=80)
    paths = synth.write_pair(p, str(tmp / f"p{seed}"))
    with contextlib.redirect_stdout(io.StringIO()):
        m1, m2, n1, n2 = cli._resample_pair(paths["A"], paths["B"], p.a.threshold, p.b.threshold, s, 8.0)
        n1 = V.fastVEC(m1, n1, 8.0)
        n2 = V.fastVEC(m2, n2, 8.0)
    return n1, n2


def test_uneven_pairs_graph_equals_eager_and_oracle(gpu, tmp_path):
    pairs = [_resampled_pair(tmp_path, 40 + i, e, a, 4.0)
             for i, (e, a) in enumerate([(60, 200), (72, 260), (60, 220), (84, 300)])]
    edges = [p[0].xdim for p in pairs]
    assert len(set(edges)) >= 2, edges
    g = batch.search_pairs(pairs, ang=30.0, topn=10, use_graph=True)
    e = batch.search_pairs(pairs, ang=30.0, topn=10, use_graph=False)
    for a, b in zip(g, e):
        assert np.array_equal(a["scores"], b["scores"]) and np.array_equal(a["trans"], b["trans"])
        assert a["top"].tolist() == b["top"].tolist()
    # one pair against the oracle's exhaustive search (no refinement)
    t, s = pairs[1]
    grid = lambda m: ref.MapGrid(data=np.asarray(m.data, np.float64), xwidth=np.float32(m.xwidth), orig=m.orig,
                                 vec=m.vec, cent=m.cent, xdim=m.xdim, ave=m.ave, std=m.std,
                                 std_norm_ave=m.std_norm_ave)
    o = ref.search_map_fft(grid(t), grid(s), TopN=10, ang=30.0, prec="ref32", refine=False)
    assert g[1]["top"].tolist() == list(o["topn"])
    assert np.allclose(g[1]["scores"], o["scores"], rtol=1e-4, atol=1e-6 * np.abs(o["scores"]).max())


def test_two_streams_identical_and_capturable(gpu):
    """Two batches in flight (second workspace on a forked stream): results
    bit-identical to one stream, eagerly and in graph replays."""
    n = 26
    rng = np.random.default_rng(31)
    td, tv = _maps(rng, n)
    sd, sv = _maps(rng, n)
    rot = torch.from_numpy(rotation_rows(rng.uniform(0, 360, (53, 3)))).cuda()
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=8)
    plan.set_target(engine.to_dev(td, torch.float64), engine.to_dev(tv, torch.float32))
    plan.set_search(engine.to_dev(sd, torch.float64), engine.to_dev(sv, torch.float32))
    print("[case] streams 1", flush=True)
    b1, t1 = plan.score_rows(rot)
    b1, t1 = b1.cpu().numpy(), t1.cpu().numpy()
    plan.tune(streams=2)
    print("[case] streams 2 eager", flush=True)
    b2, t2 = plan.score_rows(rot)
    torch.cuda.synchronize()
    assert np.array_equal(b2.cpu().numpy(), b1) and np.array_equal(t2.cpu().numpy(), t1)
    gs = batch.GraphedSearch(plan, rot, 10)
    for rep in range(3):
        print("[case] streams 2 replay", rep, flush=True)
        b, t, _ = gs.replay()
        torch.cuda.synchronize()
        assert np.array_equal(b.cpu().numpy(), b1) and np.array_equal(t.cpu().numpy(), t1)
        plan.score_rows(rot[:16])
        torch.cuda.synchronize()


def test_smaller_support_after_larger_clears_planes(gpu):
    """A second query with a smaller support box on the same plan: the planes
    written for the first box are cleared (results equal a fresh plan's)."""
    n = 30
    rng = np.random.default_rng(41)
    td, tv = _maps(rng, n)
    big_d, big_v = _maps(rng, n, 0.4)
    small_d, small_v = _maps(rng, n, 0.4)
    g = np.indices((n, n, n)).transpose(1, 2, 3, 0) - n / 2
    small = np.linalg.norm(g, axis=-1) > n / 6
    small_d[small] = 0
    small_v[small] = 0
    rot = torch.from_numpy(rotation_rows(rng.uniform(0, 360, (12, 3)))).cuda()
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=4)
    plan.set_target(engine.to_dev(td, torch.float64), engine.to_dev(tv, torch.float32))
    plan.set_search(engine.to_dev(big_d, torch.float64), engine.to_dev(big_v, torch.float32))
    plan.score_rows(rot)
    plan.set_search(engine.to_dev(small_d, torch.float64), engine.to_dev(small_v, torch.float32))
    b, t = plan.score_rows(rot)
    fresh = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=4)
    fresh.set_target(engine.to_dev(td, torch.float64), engine.to_dev(tv, torch.float32))
    fresh.set_search(engine.to_dev(small_d, torch.float64), engine.to_dev(small_v, torch.float32))
    fb, ft = fresh.score_rows(rot)
    assert plan.info()["box_width"] < n
    assert np.array_equal(b.cpu().numpy(), fb.cpu().numpy()) and np.array_equal(t.cpu().numpy(), ft.cpu().numpy())
