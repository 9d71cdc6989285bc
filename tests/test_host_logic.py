"""Host-side logic of the package on the CPU: the Euler-grid indexing the
proposer's neighbour sweep and coarse-row ranking rely on (brute force over
the grid), the bench's one-GPU contract and byte model, the traffic line it
reads from the committed profile, the batch and autotuner choices, and the
compile-time FFT plan list against the descriptor's stage order."""
import numpy as np
import pytest

from vesper_python_amd import proposer
from vesper_python_amd.geometry import calc_angle_comb


@pytest.mark.parametrize("ang", [30.0, 10.0])
def test_grid_shape_and_neighbours_brute_force(ang):
    a = calc_angle_comb(ang)
    nx, nz = proposer.grid_shape(a)
    assert nx * nx * nz == len(a)
    rng = np.random.default_rng(int(ang))
    for r in rng.choice(len(a), 12, replace=False).tolist() + [0, len(a) - 1]:
        d = np.abs(a - a[r])
        d[:, :2] = np.minimum(d[:, :2], 360 - d[:, :2])
        exp = np.flatnonzero((d <= ang + 1e-9).all(axis=1))
        assert np.array_equal(proposer.neighbour_rows([r], nx, nz), exp)


def test_grid_shape_rejects_other_row_sets():
    a = calc_angle_comb(30.0)
    assert proposer.grid_shape(a[::-1]) is None
    assert proposer.grid_shape(a[:-1]) is None


def test_nearest_coarse_rows_brute_force():
    """The coarse and hybrid proposal policies rank a fine row by its nearest
    row of the coarse grid: each returned row is a calc_angle_comb row of the
    coarse step, each of its angles is at the smallest distance any coarse
    value of that axis reaches (x and y wrap at 360, z does not; midpoints
    may go either way), and the distance returned is their Euclidean norm."""
    fine, coarse = calc_angle_comb(5.0), calc_angle_comb(30.0)
    rng = np.random.default_rng(3)
    pick = fine[rng.choice(len(fine), 400, replace=False)]
    rows, dist = proposer.nearest_coarse_rows(pick, 30.0)
    got = coarse[rows]
    wrap = lambda d: np.abs((d + 180.0) % 360.0 - 180.0)
    axes = (np.arange(0, 360, 30.0), np.arange(0, 360, 30.0), np.arange(0, 180.0 + 1e-9, 30.0))
    dev = np.empty_like(pick)
    for ax, vals in enumerate(axes):
        d_all = pick[:, ax, None] - vals[None, :]
        d_all = wrap(d_all) if ax < 2 else np.abs(d_all)
        d_got = pick[:, ax] - got[:, ax]
        d_got = wrap(d_got) if ax < 2 else np.abs(d_got)
        assert np.allclose(d_got, d_all.min(axis=1)), ax
        dev[:, ax] = d_got
    assert np.allclose(dist, np.linalg.norm(dev, axis=1))


def test_propose_rounds_rejects_unknown_policy():
    assert "hybrid" in proposer.POLICIES
    with pytest.raises(ValueError):
        proposer.propose_rounds(None, None, None, policy="best")


# ------------------------------------------------------------ bench and plans

import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    """This build measures one GPU: --gpus 2 exits with a message (status 2)
    before touching a device."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "error" in line and line["n_gpus"] == 2


def test_stage_bytes_model():
    """Algorithmic bytes per launch at N = 175, box 96, 32 rows: the x pass
    reads 3 box-slab half spectra per row and the target once, writes one
    half spectrum per row."""
    sys.path.insert(0, ROOT)
    import bench
    sb = bench._stage_bytes(175, 96, 32)
    nh = 175 * 175 * 88
    assert sb["xcorr"] == 32 * (3 * 96 * 175 * 88 * 8 + nh * 8) + 3 * nh * 8 == 1_890_011_200
    assert sb["inv_yz"] == 32 * nh * 8
    assert sb["fwd_zy"] == 32 * 3 * (96 * 97 * 96 * 4 + 96 * 175 * 88 * 8)
    assert sb["gather"] == 32 * 28 * 96 ** 3


def test_pmc_traffic_line_matches_only_its_schedule():
    sys.path.insert(0, ROOT)
    import bench
    got = bench.pmc_traffic("pair200", 32, {"xpass_variant": 3, "xpass_tile": 8}, 96)
    assert got is not None and 1.8e9 < got < 2.3e9
    assert bench.pmc_traffic("pair200", 32, {"xpass_variant": 1, "xpass_tile": 8}, 96) is None
    assert bench.pmc_traffic("pair200", 16, {"xpass_variant": 3, "xpass_tile": 8}, 96) is None


def test_auto_batch_and_tuner_candidates():
    from vesper_python_amd import autotune
    from vesper_python_amd.batch import auto_batch
    assert auto_batch(175) == 32
    assert auto_batch(24) == 128 and auto_batch(400) == 8
    assert auto_batch(175, 7) < auto_batch(175, 3)
    c175 = autotune.candidates(175)
    assert dict(xpass_variant=3, xpass_tile=8) in c175
    assert any(c["xpass_variant"] == 2 for c in c175)
    assert not any(c["xpass_variant"] == 2 for c in autotune.candidates(175, groups=2))
    assert not any(c["xpass_variant"] == 2 for c in autotune.candidates(250))
    # the register x pass holds N x tile <= 2048 values
    assert all(c["xpass_variant"] != 3 or 300 * c["xpass_tile"] <= 2048 for c in autotune.candidates(300))


def _factorize(n):
    """fft_desc_create's stage order: radix 8s, 4s, 2s, then odd radices ascending."""
    r, m = [], n
    for p in (8, 4, 2):
        while m % p == 0:
            r.append(p)
            m //= p
    for p in (3, 5, 7, 11, 13):
        while m % p == 0:
            r.append(p)
            m //= p
    return r if m == 1 else None


def _fftw_fast(n):
    m, e = n, 0
    for p in (2, 3, 5, 7):
        while m % p == 0:
            m //= p
    for p in (11, 13):
        while m % p == 0:
            m //= p
            e += 1
    return m == 1 and e <= 1


@pytest.mark.parametrize("name,lo,hi", [("StaticLengths", 90, 180), ("SmallLengths", 16, 89),
                                         ("MediumLengths", 181, 200), ("TileLengths", 201, 256),
                                         ("LargeLengths", 257, 400)])
def test_static_lengths_follow_the_descriptor_order(name, lo, hi):
    """Every compile-time plan lists its radices in fft_desc_create's order
    (else the host check would silently fall back to the runtime plan); the
    16 x 1024-slab list is every FFTW-fast length from 90 to 180, the
    8 x 512 one every such length from 16 to 89 with two stages or more, the
    24 x 1024 one every such length from 181 to 200, the x-pass-only ones
    every such length from 201 to 256 (8-column tiles) and from 257 to 400
    (4-column tiles)."""
    src = open(os.path.join(ROOT, "vesper_python_amd", "csrc", "vx_plans.hpp")).read()
    m = re.search(r"using %s = PlanList<(.*?)>;" % name, src, re.S)
    assert m
    plans = [[int(v) for v in g.split(",")] for g in re.findall(r"StaticPlan<([0-9, ]+)>", m.group(1))]
    from vesper_python_amd.autotune import radix_orders
    lengths = []
    for radices in plans:
        n = int(np.prod(radices))
        # the default order, or another order VX_TUNE_RADIX_ORDER gives
        assert tuple(radices) in radix_orders(n).values(), (n, radices)
        if radices == _factorize(n):
            lengths.append(n)
    want = [n for n in range(lo, hi + 1) if _fftw_fast(n) and (hi > 89 or len(_factorize(n)) >= 2)]
    assert lengths == want
