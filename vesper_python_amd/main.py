"""Command line of the reference (main.py) on MI355X:

    python -m vesper_python_amd.main orig -a A.mrc -b B.mrc [-t T] [-T T] [-g G] [-s S] [-A A] [-N N]
                                          [-S S] [-M V|O|C|P|L] [-E E] [-o O]
    python -m vesper_python_amd.main prob -a A.mrc -npa A.npy -b B.mrc -npb B.npy [-alpha ALPHA] ...

Same subcommands, options, defaults and printed output as main.py:68-370.
Extra long options (no reference counterpart): --batch (rotations per
launch).
"""
from __future__ import annotations

import argparse
import copy

import numpy as np

from . import VESPER_1_prob_new as V
from .VESPER_1_prob_new import fastVEC, mrc_obj, mrc_set_vox_size, search_map_fft, search_map_fft_prob

MODE_NAMES = {"V": "VecProduct", "O": "Overlap", "C": "CC", "P": "PCC", "L": "Laplacian"}


def _resample_pair(objA, objB, threshold1, threshold2, voxel_spacing, bandwidth):
    """main.py:158-191 (and alpha_is_zero, main.py:17-50)."""
    mrc1 = mrc_obj(objA)
    mrc2 = mrc_obj(objB)
    mrc1, mrc_N1 = mrc_set_vox_size(mrc1, threshold1, voxel_spacing)
    mrc2, mrc_N2 = mrc_set_vox_size(mrc2, threshold2, voxel_spacing)
    if mrc_N1.xdim > mrc_N2.xdim:
        dim = mrc_N2.xdim = mrc_N2.ydim = mrc_N2.zdim = mrc_N1.xdim
        for i in range(3):
            mrc_N2.orig[i] = mrc_N2.cent[i] - 0.5 * voxel_spacing * mrc_N2.xdim
    else:
        dim = mrc_N1.xdim = mrc_N1.ydim = mrc_N1.zdim = mrc_N2.xdim
        for i in range(3):
            mrc_N1.orig[i] = mrc_N1.cent[i] - 0.5 * voxel_spacing * mrc_N1.xdim
    for m in (mrc_N1, mrc_N2):
        m.dens = np.zeros((dim ** 3, 1))
        m.vec = np.zeros((dim, dim, dim, 3), dtype="float32")
        m.data = np.zeros((dim, dim, dim))
    return mrc1, mrc2, mrc_N1, mrc_N2


def alpha_is_zero(objA, objB, threshold1, threshold2, bandwidth, voxel_spacing, angle_spacing, topN, showPDB,
                  modeVal, evalMode, folder=None):
    """main.py:17-65 (the reference reads ``folder`` from module scope)."""
    mrc1, mrc2, mrc_N1, mrc_N2 = _resample_pair(objA, objB, threshold1, threshold2, voxel_spacing, bandwidth)
    mrc_N1 = fastVEC(mrc1, mrc_N1, bandwidth)
    mrc_N2 = fastVEC(mrc2, mrc_N2, bandwidth)
    modeVal = MODE_NAMES.get(modeVal, modeVal)
    return search_map_fft(mrc_N1, mrc_N2, TopN=topN, ang=angle_spacing, mode=modeVal, is_eval_mode=evalMode,
                          showPDB=showPDB, folder=folder)


def resample_prob(objA, objB, npA, npB, threshold1, threshold2, voxel_spacing, bandwidth):
    """The prob command's map preparation (main.py:302-361): density maps and
    the four probability channels of each (.npy in MRC order (nz, ny, nx, 4),
    swapped like the density), each thresholded (channels at 0) and sized on
    its own dmax, all ten grids padded to the largest edge, then resampled
    and smoothed by fastVEC.  Returns (mrc_N1, mrc_N2, N1p, N2p)."""
    prob_maps = np.load(npA).astype(np.float32)
    prob_maps_chain = np.load(npB).astype(np.float32)
    mrc1, mrc2 = mrc_obj(objA), mrc_obj(objB)
    mrc1_p = [mrc_obj(objA) for _ in range(4)]
    mrc2_p = [mrc_obj(objB) for _ in range(4)]
    for k in range(4):
        mrc2_p[k].data = np.swapaxes(prob_maps_chain[:, :, :, k], 0, 2)
        mrc1_p[k].data = np.swapaxes(prob_maps[:, :, :, k], 0, 2)
    print("\n###Generating Params for Resampling Density Map 1###")
    mrc1, mrc_N1 = mrc_set_vox_size(mrc1, threshold1, voxel_spacing)
    print("\n###Generating Params for Resampling Density Map 2###")
    mrc2, mrc_N2 = mrc_set_vox_size(mrc2, threshold2, voxel_spacing)
    print("\n###Generating Params for Resampling Probability Map 1###")
    res1 = [mrc_set_vox_size(m, th=0.0, voxel_size=voxel_spacing) for m in mrc1_p]
    print("\n###Generating Params for Resampling Probability Map 2###")
    res2 = [mrc_set_vox_size(m, th=0.0, voxel_size=voxel_spacing) for m in mrc2_p]
    mrc1_p = [r[0] for r in res1]
    mrc2_p = [r[0] for r in res2]
    N1p = [r[1] for r in res1]
    N2p = [r[1] for r in res2]
    mrc_list = [mrc_N1, mrc_N2, *N1p, *N2p]
    max_dim = np.max([m.xdim for m in mrc_list])
    for m in mrc_list:
        if m.xdim != max_dim:
            m.xdim = m.ydim = m.zdim = max_dim
            m.orig = m.cent - 0.5 * voxel_spacing * m.xdim
    for m in mrc_list:
        m.dens = np.zeros((max_dim ** 3, 1), dtype="float32")
        m.data = np.zeros((max_dim, max_dim, max_dim), dtype="float32")
        m.vec = np.zeros((max_dim, max_dim, max_dim, 3), dtype="float32")
    print("\n###Processing MAP1 Resampling###")
    mrc_N1 = fastVEC(mrc1, mrc_N1, dreso=bandwidth)
    N1p = [fastVEC(s, d, dreso=bandwidth) for s, d in zip(mrc1_p, N1p)]
    print("\n###Processing MAP2 Resampling###")
    mrc_N2 = fastVEC(mrc2, mrc_N2, dreso=bandwidth)
    N2p = [fastVEC(s, d, dreso=bandwidth) for s, d in zip(mrc2_p, N2p)]
    return mrc_N1, mrc_N2, N1p, N2p


def build_parser():
    parser = argparse.ArgumentParser()
    sub = parser.add_subparsers(dest="command")
    orig = sub.add_parser("orig")
    prob = sub.add_parser("prob")
    mode_help = ("V: vector product mode (default)\nO: overlap mode\nC: Cross Correlation Coefficient Mode\n"
                 "P: Pearson Correlation Coefficient Mode\nL: Laplacian Filtering Mode")
    orig.add_argument("-a", type=str, required=True, help="MAP1.mrc (large)")
    orig.add_argument("-b", type=str, required=True, help="MAP2.mrc (small)")
    orig.add_argument("-t", type=float, help="Threshold of density map1")
    orig.add_argument("-T", type=float, help="Threshold of density map2")
    orig.add_argument("-g", type=float, default=16.0,
                      help="Bandwidth of the Gaussian filter def=16.0, sigma = 0.5*[value entered]")
    orig.add_argument("-s", type=float, default=7.0, help="Sampling voxel spacing def=7.0")
    orig.add_argument("-A", type=float, default=30.0, help="Sampling angle spacing def=30.0")
    orig.add_argument("-N", type=int, default=10, help="Refine Top [int] models def=10")
    orig.add_argument("-S", type=bool, default=False, help="Show topN models in PDB format def=false")
    orig.add_argument("-M", type=str, default="V", help=mode_help)
    orig.add_argument("-E", type=bool, default=False, help="Evaluation mode of the current position def=false")
    orig.add_argument("-o", type=str, default=None, help="Output folder name")
    orig.add_argument("--batch", type=int, default=None, help="rotations per kernel launch")

    prob.add_argument("-a", type=str, required=True, help="MAP1.mrc (large)")
    prob.add_argument("-npa", type=str, required=True, help="numpy array for Predictions for map 1")
    prob.add_argument("-b", type=str, required=True, help="MAP2.mrc (small)")
    prob.add_argument("-npb", type=str, required=True, help="numpy array for Predictions for map 2")
    prob.add_argument("-alpha", type=float, default=0.5, required=False,
                      help="The weighting parameter for alpha mixing def=0.0")
    prob.add_argument("-t", type=float, help="Threshold of density map1")
    prob.add_argument("-T", type=float, help="Threshold of density map2")
    prob.add_argument("-g", type=float, default=16.0,
                      help="Bandwidth of the Gaussian filter def=16.0, sigma = 0.5*[value entered]")
    prob.add_argument("-s", type=float, default=7.0, help="Sampling voxel spacing def=7.0")
    prob.add_argument("-A", type=float, default=30.0, help="Sampling angle spacing def=30.0")
    prob.add_argument("-N", type=int, default=10, help="Refine Top [int] models def=10")
    prob.add_argument("-S", type=bool, default=False, help="Show topN models in PDB format def=false")
    prob.add_argument("-M", type=str, default="V", help=mode_help)
    prob.add_argument("-E", type=bool, default=False, help="Evaluation mode of the current position def=false")
    prob.add_argument("-P", type=int, default=4, help="Number of processors to use def=4")
    prob.add_argument("-vav", type=float, help="Pre-computed average for density map")
    prob.add_argument("-vstd", type=float, help="Pre-computed standard deviation for density map")
    prob.add_argument("-pav", type=float, help="Pre-computed average for probability map")
    prob.add_argument("-pstd", type=float, help="Pre-computed standard deviation for probability map")
    prob.add_argument("-o", type=str, default=None, help="Output folder name")
    prob.add_argument("--batch", type=int, default=None, help="rotations per kernel launch")
    return parser


def main(argv=None):
    args = build_parser().parse_args(argv)
    if getattr(args, "batch", None):
        V.SEARCH_BATCH = int(args.batch)
    if args.command == "orig":
        folder = args.o
        objA, objB = args.a, args.b
        threshold1 = 0.00 if args.t is None else args.t
        threshold2 = 0.00 if args.T is None else args.T
        bandwidth, voxel_spacing, angle_spacing, topN = args.g, args.s, args.A, args.N
        showPDB, modeVal, evalMode = args.S, args.M, args.E
        print(objA, objB, threshold1, threshold2, bandwidth, voxel_spacing, angle_spacing, topN, showPDB, modeVal,
              evalMode)
        mrc1, mrc2, mrc_N1, mrc_N2 = _resample_pair(objA, objB, threshold1, threshold2, voxel_spacing, bandwidth)
        print()
        mrc_N1 = fastVEC(mrc1, mrc_N1, bandwidth)
        mrc_N2 = fastVEC(mrc2, mrc_N2, bandwidth)
        print()
        modeVal = MODE_NAMES.get(modeVal, modeVal)
        return search_map_fft(mrc_N1, mrc_N2, TopN=topN, ang=angle_spacing, mode=modeVal, is_eval_mode=evalMode,
                              showPDB=showPDB, folder=folder)
    if args.command == "prob":
        folder = args.o
        objA, objB = args.a, args.b
        npA, npB = args.npa, args.npb
        probA, probB = args.a, args.b
        threshold1 = 0.00 if args.t is None else args.t
        threshold2 = 0.00 if args.T is None else args.T
        alpha = args.alpha
        bandwidth, voxel_spacing, angle_spacing, topN = args.g, args.s, args.A, args.N
        vave = -10 if args.vav is None else args.vav
        vstd = -10 if args.vstd is None else args.vstd
        pave = -10 if args.pav is None else args.pav
        pstd = -10 if args.pstd is None else args.pstd
        showPDB, modeVal, evalMode = args.S, args.M, args.E
        num_processors = args.P
        print(objA, objB, probA, probB, threshold1, threshold2, bandwidth, voxel_spacing, angle_spacing, topN,
              showPDB, modeVal, evalMode, alpha, vave, vstd, pave, pstd)
        if alpha == 0.0:
            return alpha_is_zero(objA, objB, threshold1, threshold2, bandwidth, voxel_spacing, angle_spacing, topN,
                                 showPDB, modeVal, evalMode, folder)
        mrc_N1, mrc_N2, N1p, N2p = resample_prob(objA, objB, npA, npB, threshold1, threshold2, voxel_spacing,
                                                 bandwidth)
        return search_map_fft_prob(*N1p, mrc_N1, mrc_N2, *N2p, ang=angle_spacing, alpha=alpha, TopN=topN,
                                   num_proc=num_processors, vave=vave, vstd=vstd, pave=pave, pstd=pstd,
                                   showPDB=showPDB, folder=folder)
    build_parser().print_help()
    return None


if __name__ == "__main__":
    main()
