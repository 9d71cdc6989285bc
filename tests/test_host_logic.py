"""Host-side logic of the package on the CPU: the Euler-grid indexing the
proposer's neighbour sweep relies on, checked by brute force over the grid."""
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
