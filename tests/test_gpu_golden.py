"""The HIP path against the committed golden fixtures (tests/golden/):
rotations bit-exact, search scores within 1e-4 of the FP64 fixture with the
FP64 translation, get_score and fastVEC to FP64 rounding."""
import os

import numpy as np
import pytest
import torch

from vesper_python_amd import _native as nat
from vesper_python_amd import engine
from vesper_python_amd.geometry import rotation_rows

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL = 1e-4


def _load(name):
    return np.load(os.path.join(GOLD, name))


@pytest.mark.parametrize("n", [9, 10])
def test_rotate_map_matches_fixture(gpu, n):
    g = _load(f"rot_mrc_n{n}.npz")
    d = engine.to_dev(g["data"], torch.float64)
    v = engine.to_dev(g["vec"], torch.float32)
    for i, a in enumerate(g["angles"]):
        print("[case] golden rotate", n, a, flush=True)
        od, ov, _ = engine.rotate_map(rotation_rows(a)[0], d, v)
        assert np.array_equal(od.cpu().numpy(), g["out_data"][i])
        assert np.array_equal(ov.cpu().numpy(), g["out_vec"][i])


@pytest.mark.parametrize("name,mode", [("search_VecProduct_n12.npz", nat.VX_MODE_VEC),
                                       ("search_CC_n11.npz", nat.VX_MODE_CC)])
def test_score_rows_match_fixture(gpu, name, mode):
    g = _load(name)
    n = g["tdata"].shape[0]
    print("[case] golden search", name, flush=True)
    plan = engine.SearchPlan(n, mode, batch=3)
    plan.set_target(engine.to_dev(g["tdata"], torch.float64), engine.to_dev(g["tvec"], torch.float32), None,
                    float(g["ave"]))
    plan.set_search(engine.to_dev(g["sdata"], torch.float64), engine.to_dev(g["svec"], torch.float32), None,
                    float(g["ave"]))
    best, trans = plan.score_rows(torch.from_numpy(rotation_rows(g["angles"])).cuda())
    best = best[:, 0].double().cpu().numpy()
    trans = trans[:, 0].cpu().numpy()
    scale = np.abs(g["best64"]).max()
    assert np.all(np.abs(best - g["best64"]) <= RTOL * scale)
    assert np.array_equal(trans, g["trans64"])


def test_get_score_matches_fixture(gpu):
    g = _load("get_score_n10.npz")
    for i, tr in enumerate(g["trans"]):
        r = engine.get_score(engine.to_dev(g["tdata"], torch.float64), engine.to_dev(g["tvec"], torch.float32),
                             float(g["ave"]), engine.to_dev(g["sdata"], torch.float64),
                             engine.to_dev(g["svec"], torch.float32), tr)
        assert r["Nm"] == g["Nm"][i] and r["total"] == g["total"][i]
        for k in ("cc_raw", "score"):
            assert np.isclose(r[k], g[k][i], rtol=1e-12), k
        assert np.isclose(r["pcc_raw"], g["pcc_raw"][i], rtol=1e-10, atol=1e-12)


def test_fastvec_matches_fixture(gpu):
    g = _load("fastvec_n9.npz")
    d, v = engine.fastvec(g["src"], np.float32(g["src_width"]), g["src_orig"], int(g["n"]),
                          float(g["dest_width"]), g["dest_orig"], float(g["dreso"]))
    d = d.cpu().numpy()
    assert np.array_equal(d != 0, g["data"] != 0)
    assert np.allclose(d, g["data"], rtol=1e-12, atol=1e-300)
    assert np.max(np.abs(v.cpu().numpy() - g["vec"])) <= 4e-7


def test_interpolate_matches_fixture(gpu):
    from vesper_python_amd import interpolation as ip
    g = _load("interp_9x8x7.npz")
    d = torch.from_numpy(g["grid"].copy()).cuda()
    assert ip.interpolate_device(d) == tuple(int(v) for v in g["parity"])
    assert np.array_equal(d.cpu().numpy(), g["out"])
