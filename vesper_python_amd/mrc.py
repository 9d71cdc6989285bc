"""Minimal MRC2014 reader/writer with mrcfile 1.3's semantics for the fields
the reference reads (VESPER_1_prob_new.py:21-43): data in file order
(nz, ny, nx), voxel size = cella / (mx, my, mz) as float32, origin = the
header ORIGIN words as float32.  nxstart and mapc/mapr/maps are ignored, as
mrcfile's ``data`` and ``voxel_size`` ignore them.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

_MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}


@dataclass
class MrcMap:
    data: np.ndarray          # (nz, ny, nx)
    voxel_size: np.ndarray    # float32 (x, y, z)
    origin: np.ndarray        # float32 (x, y, z)
    nx: int
    ny: int
    nz: int


def read_mrc(path: str) -> MrcMap:
    with open(path, "rb") as f:
        head = f.read(1024)
        if len(head) < 1024:
            raise ValueError(f"{path}: truncated MRC header")
        machst = head[212:214]
        endian = ">" if machst[0] == 0x11 else "<"
        ints = np.frombuffer(head, dtype=endian + "i4", count=256)
        flts = np.frombuffer(head, dtype=endian + "f4", count=256)
        nx, ny, nz, mode = (int(v) for v in ints[0:4])
        mx, my, mz = (int(v) for v in ints[7:10])
        cella = flts[10:13].astype(np.float32)
        nsymbt = int(ints[23])
        origin = flts[49:52].astype(np.float32)
        if mode not in _MODES:
            raise ValueError(f"{path}: unsupported MRC mode {mode}")
        dt = np.dtype(_MODES[mode]).newbyteorder(endian)
        f.seek(1024 + nsymbt)
        count = nx * ny * nz
        data = np.frombuffer(f.read(count * dt.itemsize), dtype=dt, count=count)
    data = data.reshape(nz, ny, nx).astype(dt.newbyteorder("="), copy=True)
    # float32(cella / m) computed in float64 first, as mrcfile stores it
    vs = np.array([np.float64(c) / m if m else 0.0 for c, m in zip(cella, (mx, my, mz))],
                  dtype=np.float64).astype(np.float32)
    return MrcMap(data=data, voxel_size=vs, origin=origin, nx=nx, ny=ny, nz=nz)


def write_mrc(path: str, data_zyx: np.ndarray, voxel_size=1.0, origin=(0.0, 0.0, 0.0), start=(0, 0, 0),
              cella=None) -> None:
    """Write a float32 MRC2014 file; data given in file order (nz, ny, nx).
    `start` fills NXSTART/NYSTART/NZSTART; `cella` (x, y, z), when given,
    replaces voxel_size x (nx, ny, nz) as the cell lengths."""
    arr = np.ascontiguousarray(data_zyx, dtype="<f4")
    nz, ny, nx = arr.shape
    vs = np.broadcast_to(np.asarray(voxel_size, dtype=np.float32), (3,))
    cell = ([float(vs[0]) * nx, float(vs[1]) * ny, float(vs[2]) * nz] if cella is None
            else [float(c) for c in np.asarray(cella, dtype=np.float32)])
    h = bytearray(1024)
    struct.pack_into("<4i", h, 0, nx, ny, nz, 2)
    struct.pack_into("<3i", h, 16, int(start[0]), int(start[1]), int(start[2]))
    struct.pack_into("<3i", h, 28, nx, ny, nz)
    struct.pack_into("<3f", h, 40, *cell)
    struct.pack_into("<3f", h, 52, 90.0, 90.0, 90.0)
    struct.pack_into("<3i", h, 64, 1, 2, 3)
    struct.pack_into("<3f", h, 76, float(arr.min()), float(arr.max()), float(arr.mean()))
    struct.pack_into("<2i", h, 88, 1, 0)
    struct.pack_into("<i", h, 108, 20140)
    o = np.asarray(origin, dtype=np.float32)
    struct.pack_into("<3f", h, 196, float(o[0]), float(o[1]), float(o[2]))
    h[208:212] = b"MAP "
    h[212:216] = bytes([0x44, 0x44, 0, 0])
    struct.pack_into("<f", h, 216, float(arr.std()))
    with open(path, "wb") as f:
        f.write(bytes(h))
        f.write(arr.tobytes())
