"""Learned proposal on the GPU: the MFMA bf16 GEMM against numpy (all
transposes and epilogues), MLP gradients from the library's backward kernels
against the FP64 oracle with the same bf16 rounding points, training, the
moment descriptors, and the proposal's top-10 recall against the exhaustive
search (the reference for recall) compared with random rows of the same
budget."""
import contextlib
import io

import numpy as np
import pytest
import torch

from oracle import mlp_ref as mr
from vesper_python_amd import _native as nat
from vesper_python_amd import engine, proposer, synth
from vesper_python_amd import VESPER_1_prob_new as V
from vesper_python_amd import main as cli
from vesper_python_amd.geometry import calc_angle_comb, rotation_rows

pytestmark = pytest.mark.gpu


def _bf(a):
    return torch.from_numpy(np.asarray(a, np.float32)).cuda().to(torch.bfloat16)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_bf16_layouts_and_epilogues(gpu, ta, tb):
    rng = np.random.default_rng(10 * ta + tb)
    M, N, K = 128, 64, 48
    A = mr.bf16(rng.standard_normal((M, K)))
    B = mr.bf16(rng.standard_normal((K, N)))
    Am = A.T.copy() if ta else A
    Bm = B.T.copy() if tb else B
    ref = A @ B
    a, b = _bf(Am), _bf(Bm)
    lda, ldb = Am.shape[1], Bm.shape[1]
    print("[case] gemm", ta, tb, flush=True)
    C = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    nat.call("vx_gemm_bf16", M, N, K, engine._p(a), lda, ta, engine._p(b), ldb, tb, engine._p(C), N, 0,
             None, None, 0, engine._stream())
    assert np.allclose(C.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)
    nat.call("vx_gemm_bf16", M, N, K, engine._p(a), lda, ta, engine._p(b), ldb, tb, engine._p(C), N, 3,
             None, None, 0, engine._stream())
    assert np.allclose(C.cpu().numpy(), 2 * ref, rtol=1e-5, atol=1e-5)
    bias = torch.from_numpy(rng.standard_normal(N).astype(np.float32)).cuda()
    O = torch.zeros((M, N), dtype=torch.bfloat16, device="cuda")
    nat.call("vx_gemm_bf16", M, N, K, engine._p(a), lda, ta, engine._p(b), ldb, tb, engine._p(O), N, 1,
             engine._p(bias), None, 0, engine._stream())
    exp = mr.bf16(np.maximum(ref + bias.cpu().numpy(), 0))
    got = O.float().cpu().numpy()
    assert np.mean(got == exp) > 0.99 and np.allclose(got, exp, rtol=1e-2, atol=1e-2)
    aux = _bf(rng.standard_normal((M, N)))
    nat.call("vx_gemm_bf16", M, N, K, engine._p(a), lda, ta, engine._p(b), ldb, tb, engine._p(O), N, 2,
             None, engine._p(aux), N, engine._stream())
    exp = mr.bf16(np.where(aux.float().cpu().numpy() > 0, ref, 0))
    assert np.allclose(O.float().cpu().numpy(), exp, rtol=1e-2, atol=1e-2)


def test_mlp_gradients_match_oracle(gpu):
    rng = np.random.default_rng(3)
    M = 150
    Xh = mr.bf16(rng.standard_normal((M, proposer.IN)) * 0.5)
    y = rng.standard_normal(M).astype(np.float32)
    X = torch.zeros((proposer.pad64(M), proposer.IN), dtype=torch.bfloat16, device="cuda")
    X[:M] = _bf(Xh)
    mlp = proposer.RotationMLP(seed=1)
    with torch.no_grad():
        for bb in mlp.b:
            bb.copy_(torch.from_numpy(rng.standard_normal(proposer.H).astype(np.float32) * 0.1))
    g, loss = mlp.grads(X, torch.from_numpy(y).cuda(), M)
    W = [w.cpu().numpy().astype(np.float64) for w in mlp.W]
    b = [x.cpu().numpy().astype(np.float64) for x in mlp.b]
    eg, eloss = mr.grads(Xh, y.astype(np.float64), W, b, mlp.w4.cpu().numpy(), mlp.b4.cpu().numpy())
    assert abs(loss.item() - eloss) <= 1e-3 * eloss
    for i, (gg, ee) in enumerate(zip(g, eg)):
        gg = gg.cpu().numpy().astype(np.float64).reshape(-1)
        ee = np.asarray(ee, np.float64).reshape(-1)
        cos = gg @ ee / (np.linalg.norm(gg) * np.linalg.norm(ee) + 1e-30)
        rel = np.linalg.norm(gg - ee) / (np.linalg.norm(ee) + 1e-30)
        print("[grad]", i, "cos", cos, "rel", rel, flush=True)
        assert cos > 0.999 and rel < 3e-2, (i, cos, rel)


def test_training_fits_a_smooth_function(gpu):
    rng = np.random.default_rng(4)
    M = 4096
    Xh = rng.uniform(-1, 1, (M, proposer.IN)).astype(np.float32)
    Xh[:, 8:] = 0
    y = (np.sin(2 * Xh[:, 0]) + Xh[:, 1] * Xh[:, 2] - 0.5 * Xh[:, 3] ** 2).astype(np.float32)
    X = _bf(Xh)
    mlp = proposer.RotationMLP(seed=2)
    losses = mlp.fit(X, torch.from_numpy(y).cuda(), M, epochs=15, batch=512, lr=2e-3)
    print("[train] losses", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.2 * losses[0]


def test_training_is_repeatable(gpu):
    """The backward sums over rows run in a fixed order: gradients, trained
    weights and moment descriptors are bit-identical across repeated runs."""
    rng = np.random.default_rng(6)
    M = 1000
    Xh = rng.uniform(-1, 1, (M, proposer.IN)).astype(np.float32)
    y = (np.sin(3 * Xh[:, 0]) + Xh[:, 1]).astype(np.float32)
    X, yd = _bf(Xh), torch.from_numpy(y).cuda()
    Xp = torch.zeros((proposer.pad64(M), proposer.IN), dtype=torch.bfloat16, device="cuda")
    Xp[:M] = X
    print("[case] repeatable grads M", M, flush=True)
    mlp = proposer.RotationMLP(seed=3)
    g1, l1 = mlp.grads(Xp, yd, M)
    g2, l2 = mlp.grads(Xp, yd, M)
    assert l1.item() == l2.item()
    for a, b in zip(g1, g2):
        assert torch.equal(a, b)
    print("[case] repeatable training", flush=True)
    runs = []
    for _ in range(2):
        m = proposer.RotationMLP(seed=3)
        losses = m.fit(X, yd, M, epochs=4, batch=256, lr=2e-3, seed=1)
        runs.append((losses, [p.cpu().clone() for p in m.params]))
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1], runs[1][1]):
        assert torch.equal(a, b)
    print("[case] fit after load ignores earlier fits", flush=True)
    # load() restarts the optimiser: the fit after it is the fresh network's
    m = proposer.RotationMLP(seed=3)
    st = m.state()
    m.fit(X, yd, M, epochs=2, batch=256, lr=2e-3, seed=5)
    m.load(st)
    assert m.fit(X, yd, M, epochs=4, batch=256, lr=2e-3, seed=1) == runs[0][0]
    for a, b in zip(m.params, runs[0][1]):
        assert torch.equal(a.cpu(), b)
    n = 24
    data = rng.uniform(0, 1, (n, n, n)) * (rng.uniform(size=(n, n, n)) < 0.3)
    vec = rng.standard_normal((n, n, n, 3)).astype(np.float32) * (data > 0)[..., None]
    dd, dv = engine.to_dev(data, torch.float64), engine.to_dev(vec, torch.float32)
    assert torch.equal(proposer.descriptors(dd, dv), proposer.descriptors(dd, dv))


def test_descriptors_match_oracle(gpu):
    rng = np.random.default_rng(5)
    n = 20
    data = rng.uniform(0, 1, (n, n, n)) * (rng.uniform(size=(n, n, n)) < 0.3)
    vec = rng.standard_normal((n, n, n, 3)).astype(np.float32) * (data > 0)[..., None]
    d = proposer.descriptors(engine.to_dev(data, torch.float64), engine.to_dev(vec, torch.float32))
    assert np.allclose(d.cpu().numpy(), mr.moments(data, vec), rtol=1e-5, atol=1e-6)


def _pair(tmp, seed):
    p = synth.make_config_pair("pair100", seed)
    paths = synth.write_pair(p, str(tmp / f"p{seed}"))
    with contextlib.redirect_stdout(io.StringIO()):
        m1, m2, n1, n2 = cli._resample_pair(paths["A"], paths["B"], p.a.threshold, p.b.threshold, 7.0, 16.0)
        n1 = V.fastVEC(m1, n1, 16.0)
        n2 = V.fastVEC(m2, n2, 16.0)
    return n1, n2


def test_proposal_recall_beats_random(gpu, tmp_path):
    """Fine step -A 10 (24,624 rows, N = 24); the network is fine-tuned on the
    pair's own -A 30 landscape and FFT-scores 5 % of the fine rows."""
    fine = calc_angle_comb(10.0)
    coarse = calc_angle_comb(30.0)
    rng = np.random.default_rng(0)
    rec_mlp, rec_rand = [], []
    for seed in (5001, 5002, 5003):
        n1, n2 = _pair(tmp_path, seed)
        plan = engine.SearchPlan(n1.xdim, nat.VX_MODE_VEC, batch=16)
        plan.set_target(*n1._dev, None, n1.ave)
        plan.set_search(*n2._dev, None, n1.ave)
        rot = torch.from_numpy(rotation_rows(fine)).cuda()
        best, _ = plan.score_rows(rot)
        truth = engine.topn(best[:, 0].contiguous(), 10)
        mlp = proposer.RotationMLP(seed=seed)
        pr = proposer.Proposer(plan, *n1._dev, *n2._dev, mlp)
        pr.fine_tune(coarse, epochs=40, lr=2e-3)
        top, _, scored = pr.propose(fine, frac=0.05, topn=10)
        rec_mlp.append(proposer.recall(top, truth))
        # two rounds of 2.5 % with retraining on the rows scored in between
        top2, scored2, _ = proposer.propose_rounds(proposer.Proposer(plan, *n1._dev, *n2._dev,
                                                                     proposer.RotationMLP(seed=seed)),
                                                   coarse, fine, frac=0.05, rounds=2)
        print("[recall2]", seed, proposer.recall(top2, truth), "scored", scored2, flush=True)
        rand = rng.choice(len(fine), scored, replace=False)
        b = best[:, 0].cpu().numpy()[rand]
        rtop = rand[np.argsort(-b, kind="stable")[:10]]
        rec_rand.append(proposer.recall(rtop, truth))
        print("[recall]", seed, "mlp", rec_mlp[-1], "random", rec_rand[-1], "scored", scored, flush=True)
    assert np.mean(rec_mlp) > np.mean(rec_rand)
    assert np.mean(rec_mlp) >= 0.5


def test_proposal_at_a5_against_exhaustive_search(gpu, tmp_path):
    """The learned-proposal protocol at -A 5 (191,808 rows): the network is
    trained on the pair's own exhaustive -A 10 landscape, 5 % of the -A 5 rows
    are FFT-scored in two rounds, and recall is measured against the
    exhaustive -A 5 top-10; random rows of the same budget are the baseline."""
    import time
    fine = calc_angle_comb(5.0)
    coarse = calc_angle_comb(10.0)
    assert len(fine) == 191808
    rng = np.random.default_rng(1)
    rec, rec_rand, speed = [], [], []
    for seed in (5001, 5003):
        n1, n2 = _pair(tmp_path, seed)
        plan = engine.SearchPlan(n1.xdim, nat.VX_MODE_VEC, batch=16)
        plan.set_target(*n1._dev, None, n1.ave)
        plan.set_search(*n2._dev, None, n1.ave)
        rot = torch.from_numpy(rotation_rows(fine)).cuda()
        plan.score_rows(rot[:64])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        best, _ = plan.score_rows(rot)
        truth = engine.topn(best[:, 0].contiguous(), 10)
        torch.cuda.synchronize()
        t_exh = time.perf_counter() - t0
        t0 = time.perf_counter()
        top, scored, ncoarse = proposer.propose_rounds(
            proposer.Proposer(plan, *n1._dev, *n2._dev, proposer.RotationMLP(seed=seed)), coarse, fine,
            frac=0.05, rounds=2)
        torch.cuda.synchronize()
        t_prop = time.perf_counter() - t0
        rec.append(proposer.recall(top, truth))
        rand = rng.choice(len(fine), scored, replace=False)
        b = best[:, 0].cpu().numpy()[rand]
        rec_rand.append(proposer.recall(rand[np.argsort(-b, kind="stable")[:10]], truth))
        speed.append(t_exh / t_prop)
        print("[a5]", seed, "recall", rec[-1], "random", rec_rand[-1], "scored", scored, "+", ncoarse,
              "exhaustive s", round(t_exh, 3), "proposal s", round(t_prop, 3), flush=True)
    assert np.mean(rec) > np.mean(rec_rand)
    assert np.mean(rec) >= 0.5


def test_pretraining_across_pairs_then_held_out_pairs(gpu, tmp_path):
    """The network trained across pairs on their exhaustive -A 10 landscapes
    (the protocol's training set, scaled down to 6 pairs) proposes rows of
    held-out pairs: pretrained alone, and pretrained then refined on the
    held-out pair's own -A 30 landscape in the proposal rounds; top-10 recall
    against each held-out pair's exhaustive search, random rows as baseline."""
    fine = calc_angle_comb(10.0)
    coarse = calc_angle_comb(30.0)
    train = []
    for seed in (5101, 5102, 5103, 5104, 5105, 5106):
        n1, n2 = _pair(tmp_path, seed)
        plan = engine.SearchPlan(n1.xdim, nat.VX_MODE_VEC, batch=16)
        plan.set_target(*n1._dev, None, n1.ave)
        plan.set_search(*n2._dev, None, n1.ave)
        train.append(proposer.Proposer(plan, *n1._dev, *n2._dev, None))
    mlp = proposer.RotationMLP(seed=11)
    losses = proposer.pretrain(mlp, train, fine, epochs=8, lr=2e-3)
    print("[pretrain] loss", losses[0], "->", losses[-1], flush=True)
    assert losses[-1] < losses[0]
    st = mlp.state()
    rng = np.random.default_rng(2)
    r_pre, r_ft, r_rand = [], [], []
    for seed in (5001, 5003):
        n1, n2 = _pair(tmp_path, seed)
        plan = engine.SearchPlan(n1.xdim, nat.VX_MODE_VEC, batch=16)
        plan.set_target(*n1._dev, None, n1.ave)
        plan.set_search(*n2._dev, None, n1.ave)
        best, _ = plan.score_rows(torch.from_numpy(rotation_rows(fine)).cuda())
        truth = engine.topn(best[:, 0].contiguous(), 10)
        m = proposer.RotationMLP(seed=12)
        m.load(st)
        top, _, scored = proposer.Proposer(plan, *n1._dev, *n2._dev, m).propose(fine, frac=0.05)
        r_pre.append(proposer.recall(top, truth))
        m2 = proposer.RotationMLP(seed=13)
        m2.load(st)
        top2, scored2, _ = proposer.propose_rounds(proposer.Proposer(plan, *n1._dev, *n2._dev, m2), coarse, fine,
                                                   frac=0.05, rounds=2, epochs=10)
        r_ft.append(proposer.recall(top2, truth))
        rand = rng.choice(len(fine), scored, replace=False)
        b = best[:, 0].cpu().numpy()[rand]
        r_rand.append(proposer.recall(rand[np.argsort(-b, kind="stable")[:10]], truth))
        print("[held-out]", seed, "pretrained", r_pre[-1], "pretrained+rounds", r_ft[-1], "random", r_rand[-1],
              flush=True)
    assert np.mean(r_ft) > np.mean(r_rand)
    assert np.mean(r_ft) >= 0.5


def test_ablation_network_against_coarse_and_random(gpu, tmp_path):
    """What the trained network gains over the same search without it: on
    held-out pairs, the same budget (the -A 30 landscape, a fraction of the
    -A 10 rows in two shares, three neighbour sweeps) spent by the network
    (pretrained across pairs, refined on the pair's coarse landscape), by the
    coarse landscape alone (each fine row ranked by its nearest coarse row's
    score) and by random rows; top-10 recall against the exhaustive search."""
    fine = calc_angle_comb(10.0)
    coarse = calc_angle_comb(30.0)
    train = []
    for seed in (5101, 5102, 5103, 5104, 5105, 5106):
        n1, n2 = _pair(tmp_path, seed)
        plan = engine.SearchPlan(n1.xdim, nat.VX_MODE_VEC, batch=16)
        plan.set_target(*n1._dev, None, n1.ave)
        plan.set_search(*n2._dev, None, n1.ave)
        train.append(proposer.Proposer(plan, *n1._dev, *n2._dev, None))
    mlp = proposer.RotationMLP(seed=11)
    proposer.pretrain(mlp, train, fine, epochs=8, lr=2e-3)
    st = mlp.state()
    rec = {(p, f): [] for p in proposer.POLICIES for f in (0.005, 0.02)}
    for seed in (5001, 5002, 5003, 5004):
        n1, n2 = _pair(tmp_path, seed)
        plan = engine.SearchPlan(n1.xdim, nat.VX_MODE_VEC, batch=16)
        plan.set_target(*n1._dev, None, n1.ave)
        plan.set_search(*n2._dev, None, n1.ave)
        best, _ = plan.score_rows(torch.from_numpy(rotation_rows(fine)).cuda())
        truth = engine.topn(best[:, 0].contiguous(), 10)
        pr = proposer.Proposer(plan, *n1._dev, *n2._dev, proposer.RotationMLP(seed=12))
        for (pol, frac), out in rec.items():
            pr.mlp.load(st)
            top, _, _ = proposer.propose_rounds(pr, coarse, fine, frac=frac, rounds=2, epochs=10, policy=pol,
                                                coarse_step=30.0, seed=seed)
            out.append(proposer.recall(top, truth))
        print("[ablation]", seed, {f"{p}@{f}": v[-1] for (p, f), v in rec.items()}, flush=True)
    mean = {k: float(np.mean(v)) for k, v in rec.items()}
    print("[ablation] mean recall", mean, flush=True)
    # training is deterministic (fixed-order sums): on these pairs the
    # network's recall is at least the coarse landscape's and above random
    # rows' at both budgets
    for frac in (0.005, 0.02):
        assert mean[("mlp", frac)] >= mean[("coarse", frac)]
        assert mean[("mlp", frac)] > mean[("random", frac)]
