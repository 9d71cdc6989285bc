"""MRC reading with mrcfile 1.3's semantics for the fields the reference
uses (VESPER_1_prob_new.py:21-43): data in file order (nz, ny, nx) in the
file's mode, voxel size = float32(cella / (mx, my, mz)), origin = the header
ORIGIN words, nxstart and mapc/mapr/maps ignored.  Fixtures are built here
byte by byte with struct: both byte orders, an extended header, modes 0, 1,
2, 6 and 12, and a sampling grid (mx) different from the data size (nx)."""
import struct

import numpy as np
import pytest

from vesper_python_amd.mrc import read_mrc


def _mrc_bytes(data_zyx, mode, endian="<", cella=(10.0, 20.0, 30.0), mxyz=None, origin=(1.5, -2.25, 3.0),
               start=(0, 0, 0), mapcrs=(1, 2, 3), ext=b""):
    nz, ny, nx = data_zyx.shape
    mx, my, mz = mxyz if mxyz is not None else (nx, ny, nz)
    h = bytearray(1024)
    struct.pack_into(endian + "4i", h, 0, nx, ny, nz, mode)
    struct.pack_into(endian + "3i", h, 16, *start)
    struct.pack_into(endian + "3i", h, 28, mx, my, mz)
    struct.pack_into(endian + "3f", h, 40, *cella)
    struct.pack_into(endian + "3f", h, 52, 90.0, 90.0, 90.0)
    struct.pack_into(endian + "3i", h, 64, *mapcrs)
    struct.pack_into(endian + "i", h, 92, len(ext))           # NSYMBT
    struct.pack_into(endian + "i", h, 108, 20140)
    struct.pack_into(endian + "3f", h, 196, *origin)
    h[208:212] = b"MAP "
    h[212:216] = bytes([0x11, 0x11, 0, 0]) if endian == ">" else bytes([0x44, 0x44, 0, 0])
    dt = {0: "i1", 1: "i2", 2: "f4", 6: "u2", 12: "f2"}[mode]
    return bytes(h) + ext + np.ascontiguousarray(data_zyx).astype(endian + dt).tobytes()


@pytest.mark.parametrize("mode,dtype", [(0, np.int8), (1, np.int16), (2, np.float32), (6, np.uint16),
                                        (12, np.float16)])
@pytest.mark.parametrize("endian", ["<", ">"])
def test_modes_and_byte_orders(tmp_path, mode, dtype, endian):
    rng = np.random.default_rng(mode)
    lo, hi = (0, 200) if mode == 6 else (-100, 100)
    d = rng.integers(lo, hi, (3, 4, 5)).astype(dtype)
    p = tmp_path / f"m{mode}{'be' if endian == '>' else 'le'}.mrc"
    p.write_bytes(_mrc_bytes(d, mode, endian))
    m = read_mrc(str(p))
    assert m.data.dtype == np.dtype(dtype) and m.data.shape == (3, 4, 5)
    assert np.array_equal(m.data, d)
    assert (m.nx, m.ny, m.nz) == (5, 4, 3)
    assert m.voxel_size.dtype == np.float32
    assert np.array_equal(m.voxel_size, np.float32([10.0 / 5, 20.0 / 4, 30.0 / 3]))
    assert np.array_equal(m.origin, np.float32([1.5, -2.25, 3.0]))


def test_extended_header_sampling_grid_and_ignored_fields(tmp_path):
    """nsymbt bytes are skipped; voxel size divides cella by mx..mz, not by
    nx..nz; nxstart and the axis map do not move or permute the data."""
    d = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    ext = bytes(range(256)) * 3                                # 768 bytes of extended header
    p = tmp_path / "ext.mrc"
    p.write_bytes(_mrc_bytes(d, 2, ">", cella=(7.0, 9.0, 11.0), mxyz=(3, 6, 5), start=(-4, 2, 9),
                             mapcrs=(3, 1, 2), ext=ext))
    m = read_mrc(str(p))
    assert np.array_equal(m.data, d)
    want = np.array([np.float64(np.float32(7.0)) / 3, np.float64(np.float32(9.0)) / 6,
                     np.float64(np.float32(11.0)) / 5]).astype(np.float32)
    assert np.array_equal(m.voxel_size, want)
    assert m.voxel_size[0] != np.float32(7.0 / 4)


def test_mrc_obj_attributes(tmp_path):
    """mrc_obj (VESPER_1_prob_new.py:21-56): dims from the header, widths from
    the voxel size, cent = dims / 2, orig the header origin, data swapped to
    [x][y][z], dens the flattened file-order data."""
    from vesper_python_amd.VESPER_1_prob_new import mrc_obj
    d = np.random.default_rng(3).standard_normal((4, 5, 6)).astype(np.float32)     # (nz, ny, nx)
    p = tmp_path / "o.mrc"
    p.write_bytes(_mrc_bytes(d, 2, "<", cella=(12.0, 10.0, 8.0), origin=(0.5, 1.0, -1.5)))
    m = mrc_obj(str(p))
    assert (m.xdim, m.ydim, m.zdim) == (6, 5, 4)
    assert (m.xwidth, m.ywidth, m.zwidth) == (np.float32(2.0), np.float32(2.0), np.float32(2.0))
    assert np.array_equal(m.cent, [3.0, 2.5, 2.0])
    assert np.array_equal(m.orig, [0.5, 1.0, -1.5]) and m.orig.dtype == np.float32
    assert m.data.shape == (6, 5, 4) and m.data[5, 1, 2] == d[2, 1, 5]
    assert np.array_equal(m.dens, d.flatten())
    assert m.vec.shape == (6, 5, 4, 3) and m.vec.dtype == np.float32


def test_truncated_header_is_an_error(tmp_path):
    p = tmp_path / "t.mrc"
    p.write_bytes(b"\0" * 100)
    with pytest.raises(ValueError):
        read_mrc(str(p))
