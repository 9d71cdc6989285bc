"""Pins of the interpolation oracle (oracle/interp_ref.py): the vectorised
form equals the element-by-element restatement of the reference's loops on
small grids of every parity, both dtypes; a constant field stays constant;
known values on a hand-built grid."""
import numpy as np
import pytest

from oracle import interp_ref as ir


def _lattice(rng, shape, parity, dtype, nl=4):
    d = np.zeros(shape + (nl,), dtype)
    sl = tuple(slice(parity[a], shape[a], 2) for a in range(3))
    d[sl] = rng.uniform(0.01, 1.0, d[sl].shape).astype(dtype)
    return d


@pytest.mark.parametrize("shape", [(6, 7, 5), (5, 5, 5), (8, 6, 9), (3, 4, 2)])
@pytest.mark.parametrize("parity", [(0, 0, 0), (1, 0, 1), (1, 1, 1), (0, 1, 0)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_vectorised_equals_loops(shape, parity, dtype):
    rng = np.random.default_rng(sum(shape) + 7 * sum(parity))
    d = _lattice(rng, shape, parity, dtype)
    d[tuple(rng.integers(0, s) for s in shape) + (2,)] = 0.5    # off-lattice non-zero: kept
    a, b = d.copy(), d.copy()
    ra = ir.interpolate_loops(a)
    rb = ir.interpolate(b)
    assert ra == rb
    assert np.array_equal(a, b)


def test_constant_field_stays_constant():
    """Away from the last two planes of each axis (whose lattice points the
    reference's ranges, ending before dim - 1, leave unfilled) a constant
    lattice fills to the same constant."""
    d = np.zeros((7, 6, 5, 2), np.float32)
    d[::2, ::2, ::2] = 0.25
    ir.interpolate(d)
    assert np.all(d[:-2, :-2, :-2] == 0.25)
    assert d[0, 1, 3, 0] == np.float32((0.25 * 3 + 0.0) / 4)   # its +z neighbour was never filled


def test_known_values():
    d = np.zeros((4, 3, 3, 1), np.float64)
    d[0, 0, 0, 0] = 1.0
    d[2, 0, 0, 0] = 3.0
    ref = ir.interpolate_loops(d)
    assert ref == (0, 0, 0)
    assert d[1, 0, 0, 0] == 2.0                    # pass x: (1 + 3) / 2
    # x index 1 at z = 0, y = 0 only (j, k ranges stop before the last plane)
    e = np.zeros((3, 3, 3, 1))
    e[1, 1, 1, 0] = 4.0                            # parity (1, 1, 1); x pass starts at 0
    ir.interpolate_loops(e)
    assert e[0, 1, 1, 0] == (4.0 + 0.0) / 2        # +x neighbour 4, -x wraps to x = 2 (0)
