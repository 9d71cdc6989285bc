"""Prediction-grid interpolation and per-class MRC export
(interpolation.py:1-74, interpolation_functions.py:6-177): the step that
fills the parity-lattice holes of per-voxel class predictions before they
become prob-mode channels.  The fill runs on the GPU (vx_interpolate); the
MRC export is host file I/O.

    python -m vesper_python_amd.interpolation insert -f pred.npy -s out
    python -m vesper_python_amd.interpolation graph -f pred.npy -s out -m map.mrc
"""
from __future__ import annotations

import argparse
import ctypes
import sys

import numpy as np
import torch

from . import _native as nat
from . import engine
from .mrc import read_mrc, write_mrc


def interpolate_device(t: torch.Tensor) -> tuple:
    """In-place fill of a device grid [sx, sy, sz, nl] (float32 or float64);
    returns the reference parity (x, y, z)."""
    if t.dim() != 4 or t.dtype not in (torch.float32, torch.float64) or not t.is_contiguous():
        raise ValueError("expected a contiguous float32/float64 tensor [sx, sy, sz, nl]")
    ref = np.zeros(3, dtype=np.int32)
    sx, sy, sz, nl = (int(v) for v in t.shape)
    nat.call("vx_interpolate", engine._p(t), int(t.dtype == torch.float64), sx, sy, sz, nl,
             ref.ctypes.data_as(ctypes.c_void_p), engine._stream())
    return tuple(int(v) for v in ref)


def interpolate(data):
    """interpolate (interpolation_functions.py:6-135) on a NumPy grid
    [sx, sy, sz, nl], in place, as the reference's Numba function."""
    arr = np.asarray(data)
    if arr.dtype not in (np.float32, np.float64):
        raise TypeError(f"prediction grid must be float32 or float64, not {arr.dtype}")
    dev = engine.to_dev(arr, torch.float64 if arr.dtype == np.float64 else torch.float32)
    interpolate_device(dev)
    data[...] = dev.cpu().numpy()


def conv_prediction_to_mrc(arr, reffile, keyidx, outmapfile):
    """Write class `keyidx` of the grid [nx, ny, nz, nl] as an MRC map
    (interpolation_functions.py:137-177): data[k, j, i] = arr[i, j, k, keyidx]
    for i < nx - 2, j < ny - 2, k < nz - 2 (the rest zero), the reference
    map's origin and cell lengths, NXSTART..NZSTART = (nx, ny, nz)."""
    ref = read_mrc(reffile)
    with open(reffile, "rb") as f:
        head = f.read(1024)
    endian = ">" if head[212] == 0x11 else "<"
    cella = np.frombuffer(head, dtype=endian + "f4", count=256)[10:13].astype(np.float32)
    nx, ny, nz = arr.shape[0], arr.shape[1], arr.shape[2]
    out = np.zeros((nz, ny, nx), dtype=np.float32)
    out[:nz - 2, :ny - 2, :nx - 2] = np.transpose(np.asarray(arr)[:nx - 2, :ny - 2, :nz - 2, keyidx], (2, 1, 0))
    print(nx, ny, nz)
    write_mrc(outmapfile, out, origin=ref.origin, start=(nx, ny, nz), cella=cella)
    new = read_mrc(outmapfile)
    print("original", ref.voxel_size)
    print()
    print("new", new.voxel_size)
    print()


def main(argv=None):
    print('Usage Modes:\n')
    print('1- Insert: only interpolates the map given.')
    print('2- Graph: interpolates the map given and saves each secondary structure prediction map individually.')
    parser = argparse.ArgumentParser()
    sub = parser.add_subparsers(dest="command")
    insert = sub.add_parser("insert")
    graph = sub.add_parser("graph")
    insert.add_argument("-f", type=str, required=True, help="Map array (.npy)")
    insert.add_argument("-s", type=str, required=True, help="Name for new map to be saved (string)")
    graph.add_argument("-f", type=str, required=True, help="Map array (.npy)")
    graph.add_argument("-s", type=str, required=True, help="Name for individual maps (string)")
    graph.add_argument("-m", type=str, required=True,
                       help="Name of density map file. NOTE: VOXEL SPACIING MUST BE 1. (.mrc)")
    args = parser.parse_args(argv)
    if args.command not in ("insert", "graph"):
        return None
    orig_data = np.load(args.f)
    np.set_printoptions(precision=3)
    for a in range(4):
        print(orig_data.shape[a])
    interpolate(orig_data)
    np.save(args.s, orig_data)
    print("Saved interpolated map as: " + args.s + ".npy")
    if args.command == "graph":
        for i in range(4):
            conv_prediction_to_mrc(orig_data, args.m, i, args.s + f"_{i}.mrc")
    return orig_data


if __name__ == "__main__":
    main()
    sys.exit(0)
