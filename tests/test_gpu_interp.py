"""The GPU prediction-grid interpolation (vx_interpolate) against the oracle
(oracle/interp_ref.py): bit-exact on lattice grids of every parity, both
dtypes, with off-lattice values kept; the no-reference-voxel error; the
insert/graph CLI (.npy and the four per-class MRC maps)."""
import contextlib
import io
import os

import numpy as np
import pytest
import torch

from oracle import interp_ref as ir
from vesper_python_amd import _native as nat
from vesper_python_amd import interpolation as ip
from vesper_python_amd.mrc import read_mrc, write_mrc

pytestmark = pytest.mark.gpu


def _grid(rng, shape, parity, dtype, nl=4, density=0.8):
    d = np.zeros(tuple(shape) + (nl,), dtype)
    sl = tuple(slice(parity[a], shape[a], 2) for a in range(3))
    vals = rng.uniform(0.01, 1.0, d[sl].shape) * (rng.uniform(size=d[sl].shape) < density)
    d[sl] = vals.astype(dtype)
    first = tuple(parity)                     # make the reference voxel the lattice origin
    d[first + (0,)] = 0.5
    return d


@pytest.mark.parametrize("shape", [(9, 8, 7), (16, 16, 16), (33, 20, 41), (5, 64, 6)])
@pytest.mark.parametrize("parity", [(0, 0, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_interpolate_matches_oracle(gpu, shape, parity, dtype):
    rng = np.random.default_rng(hash((shape, parity)) % 2**32)
    d = _grid(rng, shape, parity, dtype)
    print("[case] interp", shape, parity, np.dtype(dtype).name, flush=True)
    exp = d.copy()
    ref = ir.interpolate(exp)
    dev = torch.from_numpy(d.copy()).cuda()
    got_ref = ip.interpolate_device(dev)
    assert got_ref == ref == tuple(parity)
    assert np.array_equal(dev.cpu().numpy(), exp)


def test_numpy_entry_point_in_place(gpu):
    rng = np.random.default_rng(3)
    d = _grid(rng, (12, 10, 14), (1, 0, 0), np.float32)
    exp = d.copy()
    ir.interpolate(exp)
    assert ip.interpolate(d) is None
    assert np.array_equal(d, exp)


def test_no_reference_voxel_is_an_error(gpu):
    d = torch.zeros((6, 6, 6, 4), dtype=torch.float32, device="cuda")
    d[2, 2, 2, 1] = 1.0                       # channel 0 never positive
    with pytest.raises(nat.VesperError):
        ip.interpolate_device(d)


def test_cli_insert_and_graph(gpu, tmp_path):
    rng = np.random.default_rng(4)
    d = _grid(rng, (20, 18, 16), (0, 1, 0), np.float32)
    src = str(tmp_path / "pred.npy")
    np.save(src, d)
    exp = d.copy()
    ir.interpolate(exp)
    mapf = str(tmp_path / "map.mrc")
    write_mrc(mapf, np.zeros((16, 18, 20), np.float32), voxel_size=1.0, origin=(3.0, -2.0, 5.5))
    with contextlib.redirect_stdout(io.StringIO()) as buf:
        ip.main(["graph", "-f", src, "-s", str(tmp_path / "out"), "-m", mapf])
    assert "Saved interpolated map as: " in buf.getvalue()
    assert np.array_equal(np.load(str(tmp_path / "out.npy")), exp)
    for key in range(4):
        m = read_mrc(str(tmp_path / f"out_{key}.mrc"))
        assert (m.nx, m.ny, m.nz) == (20, 18, 16)
        assert np.array_equal(m.origin, np.float32([3.0, -2.0, 5.5]))
        want = np.zeros((16, 18, 20), np.float32)
        want[:14, :16, :18] = exp[:18, :16, :14, key].transpose(2, 1, 0)
        assert np.array_equal(m.data, want)
