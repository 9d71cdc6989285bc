"""The schedule autotuner on a slab grid and on a grid that only the pencil
passes can hold: it returns a schedule the plan accepts, applies it, and the
tuned plan scores the rows exactly as the default schedule does (every FFT
schedule gives identical results; the MFMA x passes only enter within
their 1e-4 gate, checked against the FFT result)."""
import numpy as np
import pytest
import torch

from vesper_python_amd import _native as nat
from vesper_python_amd import autotune, engine
from vesper_python_amd.geometry import rotation_rows

pytestmark = pytest.mark.gpu


def _plan(n, seed, batch):
    rng = np.random.default_rng(seed)
    c = n / 2
    g = np.indices((n, n, n), dtype=np.float32).transpose(1, 2, 3, 0) - c
    r = np.linalg.norm(g, axis=-1)
    td = rng.uniform(0.1, 1.0, (n, n, n)) * (r < 0.35 * n) * (rng.uniform(size=(n, n, n)) < 0.2)
    sd = rng.uniform(0.1, 1.0, (n, n, n)) * (r < 0.2 * n) * (rng.uniform(size=(n, n, n)) < 0.5)
    tv = (rng.standard_normal((n, n, n, 3)) * (td != 0)[..., None]).astype(np.float32)
    sv = (rng.standard_normal((n, n, n, 3)) * (sd != 0)[..., None]).astype(np.float32)
    plan = engine.SearchPlan(n, nat.VX_MODE_VEC, batch=batch)
    plan.set_target(engine.to_dev(td, torch.float64), engine.to_dev(tv, torch.float32))
    plan.set_search(engine.to_dev(sd, torch.float64), engine.to_dev(sv, torch.float32))
    rot = torch.from_numpy(rotation_rows(rng.uniform(0, 360, (4 * batch, 3)))).cuda()
    return plan, rot


@pytest.mark.parametrize("n,batch", [(45, 8), (250, 4)])
def test_autotune_applies_a_valid_schedule(gpu, n, batch):
    print("[case] autotune n", n, flush=True)
    plan, rot = _plan(n, 300 + n, batch)
    b0, t0 = plan.score_rows(rot)
    b0, t0 = b0.cpu().numpy(), t0.cpu().numpy()
    rec = autotune.tune(plan, rot, sample_rows=2 * batch, use_cache=False)
    sched = rec["schedule"]
    assert "xpass_variant" in sched and sched.get("streams") in (1, 2)
    b1, t1 = plan.score_rows(rot)
    b1, t1 = b1.cpu().numpy(), t1.cpu().numpy()
    if sched["xpass_variant"] in (2, 4) or sched.get("pencil") or sched.get("radix_order"):
        # held to the autotuner's 1e-4 gate
        assert np.max(np.abs(b1 - b0)) <= 1e-4 * np.max(np.abs(b0))
    else:
        assert np.array_equal(b1, b0) and np.array_equal(t1, t0)
    plan.close()


@pytest.mark.parametrize("n,static", [(175, True), (96, True), (96, False)])
def test_radix_orders_agree(gpu, n, static):
    """Every stage order of the FFT (VX_TUNE_RADIX_ORDER) scores the rows
    within fp32 rounding of the default order: at 175 = 5 5 7 and 7 5 5, both
    compiled; at 96 = 8 4 3, 3 4 8 and 3 8 4 on compile-time or runtime plans."""
    plan, rot = _plan(n, 400 + n, 8)
    plan.tune(static_plans=static)
    ref_b, ref_t = plan.score_rows(rot)
    ref_b, ref_t = ref_b.double().cpu().numpy(), ref_t.cpu().numpy()
    orders = autotune.radix_orders(n)
    assert len(orders) >= 2
    for order, radices in orders.items():
        print("[case] radix order", n, order, radices, flush=True)
        plan.tune(radix_order=order)
        b, t = plan.score_rows(rot)
        b, t = b.double().cpu().numpy(), t.cpu().numpy()
        rel = np.max(np.abs(b - ref_b)) / np.max(np.abs(ref_b))
        print("[radix]", n, radices, "max rel", rel, "translations equal", np.mean(np.all(t == ref_t, axis=-1)),
              flush=True)
        assert rel <= 1e-5
    plan.tune(radix_order=0)
    b, t = plan.score_rows(rot)
    assert np.array_equal(b.double().cpu().numpy(), ref_b) and np.array_equal(t.cpu().numpy(), ref_t)
    plan.close()
