"""NumPy restatement of the VESPER_Python map-alignment hot path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; never by the product package.

Every function cites the reference lines it restates (paths relative to the
reference checkout).  Two dtype paths:

* ``prec="ref32"`` reproduces the reference's own dtypes under numpy-1.21
  promotion rules (numpy >= 2 changed scalar promotion and made
  ``np.fft.rfftn`` keep float32, so every cast is spelled out here):
  float32 rotation arithmetic, float32 r2c/c2r per channel, a complex128
  product against a complex128 target spectrum cast back to complex64,
  float32 channel sums.
* ``prec="ref64"`` computes the same quantities in float64 (the FP64 oracle):
  the integer gather indices are the algorithm's own (float32 positions, as
  the reference defines them) but vector rotation, FFTs, products and sums
  are float64.

Pins (see tests/test_oracle_*.py): closed forms (row enumeration, fast
lengths, refinement grid), brute-force loops on tiny inputs (circular
correlation, get_score, fastVEC), invariants (identity / 90-degree
permutations / equivalent Euler triples in the gather, single-voxel fastVEC,
translated self-copy argmax), and the notebook relation "FFT score x 1/N^3 ==
get_score DOT when nothing wraps" (Notebooks/VESPER_1.ipynb:1273 vs :1328).
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np
import scipy.fft
from scipy.spatial.transform import Rotation

MODES = ("VecProduct", "Overlap", "CC", "PCC", "Laplacian")


# --------------------------------------------------------------------------
# Grid geometry
# --------------------------------------------------------------------------

def is_fftw_fast(n: int) -> bool:
    """True iff n = 2^a 3^b 5^c 7^d 11^e 13^f with e+f <= 1.

    pyFFTW 0.12 ``next_fast_len`` (called at VESPER_1_prob_new.py:97) returns
    the smallest such length >= target (its documented rule; the reference
    notebook printed Nvox=39 and Nvox=55, Notebooks/VESPER_prob_rot.ipynb:1431,
    :1439, which a 5-smooth rule would not give)."""
    if n < 1:
        return False
    e = 0
    for p in (11, 13):
        while n % p == 0:
            n //= p
            e += 1
    if e > 1:
        return False
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def next_fast_len(target: int) -> int:
    """pyFFTW 0.12 next_fast_len (VESPER_1_prob_new.py:97)."""
    n = max(int(target), 1)
    while not is_fftw_fast(n):
        n += 1
    return n


@dataclass
class MapGrid:
    """Attributes of the reference's ``mrc_obj`` (VESPER_1_prob_new.py:18-56)
    that the hot path reads.  ``data`` is indexed [x][y][z]."""
    data: np.ndarray
    xwidth: float
    orig: np.ndarray
    vec: np.ndarray | None = None
    cent: np.ndarray | None = None
    xdim: int = 0
    ave: float | None = None
    std: float | None = None
    std_norm_ave: float | None = None
    dsum: float | None = None
    Nact: int | None = None
    extra: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.cent is None:
            nx, ny, nz = self.data.shape
            self.cent = np.array([nx * 0.5, ny * 0.5, nz * 0.5])
        if not self.xdim:
            self.xdim = int(self.data.shape[0])


def set_vox_size(m: MapGrid, th: float, voxel_size: float):
    """``mrc_set_vox_size`` (VESPER_1_prob_new.py:59-118).

    Returns (thresholded source MapGrid, resampled-grid MapGrid without data).
    ``m.xwidth`` is a float32 scalar and ``m.orig`` a float32 array as mrcfile
    hands them over (VESPER_1_prob_new.py:31-40)."""
    data = np.array(m.data, copy=True)
    if th < 0:            # :73-75 shifts only ``dens``, never ``data``
        th = 0.0
    data[data < th] = 0.0                                   # :79
    idx = np.array(np.nonzero(data)).T                      # :82
    d2 = np.linalg.norm(idx - np.asarray(m.cent), axis=1)   # :84
    dmax_vox = max(d2)                                      # :85
    dmax = float(dmax_vox) * np.float64(np.float32(m.xwidth))  # :89
    new_cent = (np.asarray(m.cent, np.float64) * np.float64(np.float32(m.xwidth))
                + np.asarray(m.orig, np.float32).astype(np.float64))  # :92
    tmp_size = 2 * dmax / voxel_size                        # :94
    n = next_fast_len(int(tmp_size))                        # :97
    new_orig = new_cent - 0.5 * n * voxel_size              # :100
    src = MapGrid(data=data, xwidth=m.xwidth, orig=m.orig, cent=m.cent)
    dst = MapGrid(data=np.zeros((n, n, n)), xwidth=voxel_size, orig=new_orig,
                  cent=new_cent, xdim=n)
    # the printed "#dmax=" value: the voxel-unit extent over the width (:88)
    dst.extra["dmax_vox"] = dmax_vox
    dst.extra["dmax_printed"] = str(dmax_vox / np.float32(m.xwidth))
    return src, dst


def unify_dims(a: MapGrid, b: MapGrid, voxel_spacing: float) -> int:
    """Pad both resampled grids to the larger N (main.py:166-178)."""
    if a.xdim > b.xdim:
        dim = b.xdim = a.xdim
        for i in range(3):
            b.orig[i] = b.cent[i] - 0.5 * voxel_spacing * b.xdim
    else:
        dim = a.xdim = b.xdim
        for i in range(3):
            a.orig[i] = a.cent[i] - 0.5 * voxel_spacing * a.xdim
    return dim


# --------------------------------------------------------------------------
# fastVEC (a)
# --------------------------------------------------------------------------

def fastvec(src: MapGrid, n: int, dest_width: float, dest_orig, dreso: float = 16.0,
            data_dtype=np.float64, prec: str = "ref32"):
    """``doFastVEC`` + ``calc`` (VESPER_1_prob_new.py:121-142, :180-257).

    Direct form: for each destination voxel with a non-zero source voxel under
    it, the Gaussian weight exp(-1.5 d^2 fsiv) over the axis-aligned window
    (zeros included), the weighted sum (-> data) and the unit vector from the
    voxel position to the weighted centroid (-> vec, float32 in ref32).
    The summation is NumPy's pairwise order, not the reference's sequential
    one (oracle/fastvec_ref.c keeps the sequential order)."""
    sd = np.asarray(src.data)
    dims = np.array(sd.shape)
    gstep = np.float64(np.float32(src.xwidth))
    fs = (dreso / gstep) * 0.5
    fs = fs ** 2
    fsiv = 1.0 / fs
    fmaxd = (dreso / gstep) * 2.0
    so = np.asarray(src.orig, np.float32).astype(np.float64)
    do = np.asarray(dest_orig, np.float64)
    data = np.zeros((n, n, n), dtype=np.float64)
    vec = np.zeros((n, n, n, 3), dtype=np.float32 if prec == "ref32" else np.float64)
    sd64 = sd.astype(np.float64)
    for x in range(n):
        for y in range(n):
            for z in range(n):
                xyz = np.array((x, y, z))
                pos = (xyz * dest_width + do - so) / gstep          # :196
                if (pos < 0).any() or (pos >= dims).any():           # :200-208
                    continue
                if sd[int(pos[0]), int(pos[1]), int(pos[2])] == 0:   # :210
                    continue
                stp = (pos - fmaxd).astype(np.int32)                 # :214
                stp[stp < 0] = 0
                endp = (pos + fmaxd + 1).astype(np.int32)            # :225
                endp = np.minimum(endp, dims)
                xs = np.arange(stp[0], endp[0])
                ys = np.arange(stp[1], endp[1])
                zs = np.arange(stp[2], endp[2])
                rx = (xs.astype(np.float64) - pos[0]) ** 2
                ry = (ys.astype(np.float64) - pos[1]) ** 2
                rz = (zs.astype(np.float64) - pos[2]) ** 2
                d2 = (rx[:, None, None] + ry[None, :, None]) + rz[None, None, :]
                v = sd64[stp[0]:endp[0], stp[1]:endp[1], stp[2]:endp[2]] * np.exp(-1.5 * d2 * fsiv)
                dtotal = v.sum()
                p2 = np.array([(v.sum(axis=(1, 2)) * xs).sum(),
                               (v.sum(axis=(0, 2)) * ys).sum(),
                               (v.sum(axis=(0, 1)) * zs).sum()])
                data[x, y, z] = dtotal                               # :237
                if dtotal == 0:
                    continue
                p2 = p2 * (1.0 / dtotal)
                tmp = p2 - pos
                dvec = np.sqrt(tmp[0] ** 2 + tmp[1] ** 2 + tmp[2] ** 2)
                if dvec == 0:
                    dvec = 1.0
                vec[x, y, z] = tmp * (1.0 / dvec)                    # :255
    return data.astype(data_dtype), vec


def map_stats(data):
    """Statistics after fastVEC (VESPER_1_prob_new.py:157-175)."""
    pos = data[data > 0]
    dsum = np.sum(data)
    nact = int(np.count_nonzero(data))
    ave = np.mean(pos)
    std = np.linalg.norm(pos)
    std_norm_ave = np.linalg.norm(pos - ave)
    return dict(dsum=dsum, Nact=nact, ave=ave, std=std, std_norm_ave=std_norm_ave)


# --------------------------------------------------------------------------
# Rotation rows and matrices
# --------------------------------------------------------------------------

def calc_angle_comb(ang_interval):
    """Euler rows, z slowest / x middle / y fastest (VESPER_1_prob_new.py:1396-1415)."""
    xa, za = [], []
    i = 0
    while i < 360:
        xa.append(i)
        i += ang_interval
    i = 0
    while i <= 180:
        za.append(i)
        i += ang_interval
    return np.array(np.meshgrid(xa, xa, za)).T.reshape(-1, 3)


def flip_matrix_f32(angle):
    """float32(flip(M)) with M = R.from_euler("ZYX", angle, degrees=True) and
    entries |m| <= 1e-15 zeroed (VESPER_1_prob_new.py:305-307, :271).
    Positions use its transpose (:310), vectors use it as is (:341)."""
    m = Rotation.from_euler("ZYX", angle, degrees=True).as_matrix()
    m[np.isclose(m, 0, atol=1e-15)] = 0
    return np.flip(m).astype(np.float32)


def _fma32(a, b, c):
    """Correctly rounded float32 fma(a, b, c) for float32 operands, using an
    exact float64 product, TwoSum and round-to-odd before the final rounding."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = np.asarray(c, np.float64)
    p = a * b                       # exact: 24-bit x 24-bit < 53 bits
    s = p + c
    bb = s - p
    e = (p - (s - bb)) + (c - bb)   # exact error of s
    bits = s.view(np.int64)
    fix = (e != 0) & ((bits & 1) == 0)
    if np.any(fix):
        s = np.where(fix, np.nextafter(s, np.where(e > 0, np.inf, -np.inf)), s)
    return s.astype(np.float32)


def dot3_f32(a, w, fma=True):
    """Row-vector (K=3) float32 product  sum_k a[:, k] * w[k]  in the order a
    BLAS sgemm microkernel uses for K=3: acc = a0*w0; acc = fma(a1, w1, acc);
    acc = fma(a2, w2, acc)  (``fma=False``: separately rounded products and
    sums, same order).  The reference delegates this to Numba -> BLAS sgemm
    (VESPER_1_prob_new.py:260-276); its exact FMA use is not observable here."""
    a = np.asarray(a, np.float32)
    w = np.asarray(w, np.float32)
    t = (a[:, 0] * w[0]).astype(np.float32)
    if fma:
        t = _fma32(a[:, 1], w[1], t)
        t = _fma32(a[:, 2], w[2], t)
    else:
        t = (t + a[:, 1] * w[1]).astype(np.float32)
        t = (t + a[:, 2] * w[2]).astype(np.float32)
    return t


# --------------------------------------------------------------------------
# Rotate-and-gather (b)
# --------------------------------------------------------------------------

def _rotated_rows(dim, angle, fma):
    """old positions (float32), new positions (float64) for every destination
    row in meshgrid order (VESPER_1_prob_new.py:296-310)."""
    new_pos = np.array(np.meshgrid(np.arange(dim), np.arange(dim), np.arange(dim))).T.reshape(-1, 3)
    cent = 0.5 * float(dim)
    new_pos = new_pos - cent
    f = flip_matrix_f32(angle)
    n32 = new_pos.astype(np.float32)
    old = np.empty((new_pos.shape[0], 3), np.float32)
    for j in range(3):
        # (n @ f.T)[:, j] = sum_k n_k * f[j, k]
        old[:, j] = dot3_f32(n32, f[j, :], fma)
    old = (old + np.float32(cent)).astype(np.float32)   # float32 array + python float
    return old, new_pos, cent, f


def _last_wins(dest_flat):
    """Indices (into a row list) that survive a sequential fancy assignment
    where later rows overwrite earlier ones at the same destination."""
    rev = dest_flat[::-1]
    _, first_in_rev = np.unique(rev, return_index=True)
    return np.sort(len(dest_flat) - 1 - first_in_rev)


def rot_mrc(data, vec, angle, fma=True, prec="ref32"):
    """``rot_mrc`` (VESPER_1_prob_new.py:279-354): inverse-rotate every
    destination voxel centre, keep in-bounds rows whose truncated source voxel
    has non-zero density, gather density and vector, rotate the vector, and
    scatter to int(int32(new_pos) + cent) (odd N folds the two centre planes;
    the later row wins)."""
    dim = vec.shape[0]
    old, new_pos, cent, f = _rotated_rows(dim, angle, fma)
    inb = ((old[:, 0] >= 0) & (old[:, 1] >= 0) & (old[:, 2] >= 0)
           & (old[:, 0] < dim) & (old[:, 1] < dim) & (old[:, 2] < dim))
    comb = np.hstack((old, new_pos))[inb].astype(np.int32)
    idx = comb[:, 0:3]
    keep = data[idx[:, 0], idx[:, 1], idx[:, 2]] != 0.0
    rows = comb[keep]
    src_vec = vec[rows[:, 0], rows[:, 1], rows[:, 2]]
    src_dens = data[rows[:, 0], rows[:, 1], rows[:, 2]]
    if prec == "ref32":
        nv = np.empty_like(src_vec, dtype=np.float32)
        for j in range(3):
            nv[:, j] = dot3_f32(src_vec, f[:, j], fma)
    else:
        nv = src_vec.astype(np.float64) @ f.astype(np.float64)
    dest = (rows[:, 3:6] + cent).astype(int)
    flat = (dest[:, 0] * dim + dest[:, 1]) * dim + dest[:, 2]
    sel = _last_wins(flat)
    out_vec = np.zeros(vec.shape, dtype=nv.dtype)
    out_data = np.zeros_like(data)
    d = dest[sel]
    out_vec[d[:, 0], d[:, 1], d[:, 2]] = nv[sel]
    out_data[d[:, 0], d[:, 1], d[:, 2]] = src_dens[sel]
    return out_vec, out_data


def rot_mrc_prob(data, vec, p1, p2, p3, p4, angle, fma=True, prec="ref32"):
    """``rot_mrc_prob`` (VESPER_1_prob_new.py:1211-1299): as rot_mrc, plus four
    probability channels each gathered with its own non-zero mask."""
    dim = vec.shape[0]
    out_vec, out_data = rot_mrc(data, vec, angle, fma, prec)
    old, new_pos, cent, _ = _rotated_rows(dim, angle, fma)
    inb = ((old[:, 0] >= 0) & (old[:, 1] >= 0) & (old[:, 2] >= 0)
           & (old[:, 0] < dim) & (old[:, 1] < dim) & (old[:, 2] < dim))
    comb = np.hstack((old, new_pos))[inb].astype(np.int32)
    idx = comb[:, 0:3]
    outs = []
    for p in (p1, p2, p3, p4):
        keep = p[idx[:, 0], idx[:, 1], idx[:, 2]] != 0.0
        rows = comb[keep]
        vals = p[rows[:, 0], rows[:, 1], rows[:, 2]]
        dest = (rows[:, 3:6] + cent).astype(int)
        flat = (dest[:, 0] * dim + dest[:, 1]) * dim + dest[:, 2]
        sel = _last_wins(flat)
        o = np.zeros_like(p)
        d = dest[sel]
        o[d[:, 0], d[:, 1], d[:, 2]] = vals[sel]
        outs.append(o)
    return (out_vec, out_data, *outs)


# --------------------------------------------------------------------------
# Channels, spectra, correlation (c)
# --------------------------------------------------------------------------

def laplacian_filter(arr):
    """6-neighbour stencil where arr > 0, zero outside the grid
    (VESPER_1_prob_new.py:494-526), in the reference's addition order."""
    a = np.asarray(arr)
    out = np.zeros_like(a)
    pad = np.zeros(tuple(s + 2 for s in a.shape), dtype=a.dtype)
    pad[1:-1, 1:-1, 1:-1] = a
    acc = -6.0 * a
    for sl in ((slice(2, None), slice(1, -1), slice(1, -1)),
               (slice(0, -2), slice(1, -1), slice(1, -1)),
               (slice(1, -1), slice(2, None), slice(1, -1)),
               (slice(1, -1), slice(0, -2), slice(1, -1)),
               (slice(1, -1), slice(1, -1), slice(2, None)),
               (slice(1, -1), slice(1, -1), slice(0, -2))):
        acc = acc + pad[sl]
    m = a > 0
    out[m] = acc[m]
    return out


def target_channels(t: MapGrid, mode):
    """Target channels per mode (VESPER_1_prob_new.py:601-615)."""
    if mode == "VecProduct":
        return [t.vec[..., 0], t.vec[..., 1], t.vec[..., 2]]
    d = t.data
    if mode == "Overlap":
        return [np.where(d > 0, 1.0, 0.0)]
    if mode == "CC":
        return [np.where(d > 0, d, 0.0)]
    if mode == "PCC":
        return [np.where(d > 0, d - t.ave, 0.0)]
    if mode == "Laplacian":
        return [laplacian_filter(d)]
    raise ValueError(mode)


def query_channels(new_data, new_vec, mode, ave=None):
    """Query channels per mode, cast to float32 by ``np.copyto`` into the
    float32 FFTW input (VESPER_1_prob_new.py:546-558, :1434-1438)."""
    if mode == "VecProduct":
        return [new_vec[..., 0], new_vec[..., 1], new_vec[..., 2]]
    d = new_data
    if mode == "Overlap":
        x = np.where(d > 0, 1.0, 0.0)
    elif mode == "CC":
        x = np.where(d > 0, d, 0.0)
    elif mode == "PCC":
        x = np.where(d > 0, d - ave, 0.0)
    elif mode == "Laplacian":
        x = laplacian_filter(d)
    else:
        raise ValueError(mode)
    return [x]


def target_spectra(chans):
    """conj(rfftn(x)) in float64/complex128, as numpy 1.21's pocketfft did
    (VESPER_1_prob_new.py:621-628, :846-860)."""
    return [np.conj(np.fft.rfftn(np.asarray(c, np.float64))) for c in chans]


def fft_correlate(tspecs, qchans, prec="ref32"):
    """``fft_search_best_dot`` / ``fft_search_score_trans_1d``
    (VESPER_1_prob_new.py:464-491, :557-563): per channel r2c, times the
    conjugated target spectrum, unnormalised c2r.
    corr[t] = N^3 * sum_v T[v] Q[(v + t) mod N]."""
    out = []
    for ts, q in zip(tspecs, qchans):
        shape = q.shape
        if prec == "ref32":
            qs = scipy.fft.rfftn(np.asarray(q, np.float32))          # complex64
            prod = (ts * qs.astype(np.complex128)).astype(np.complex64)
            out.append(scipy.fft.irfftn(prod, s=shape, norm="forward").astype(np.float32))
        else:
            qs = np.fft.rfftn(np.asarray(q, np.float64))
            out.append(np.fft.irfftn(ts * qs, s=shape, axes=(0, 1, 2), norm="forward"))
    return out


def find_best_trans_list(arrs):
    """Channel sum, max and first argmax in C order
    (VESPER_1_prob_new.py:357-374)."""
    s = np.zeros_like(arrs[0])
    for a in arrs:
        s = s + a
    best = np.amax(s)
    trans = tuple(int(v) for v in np.unravel_index(s.argmax(), s.shape))
    return best, trans


def find_best_trans_mixed(vres, pres, alpha, vstd, vave, pstd, pave, prec="ref32"):
    """Alpha mix of z-normalised channel sums (VESPER_1_prob_new.py:1530-1545).
    ref32: float32 array arithmetic with the float64 scalars cast to float32
    (numpy-1.21 value-based casting)."""
    sv = 0
    for a in vres:
        sv = sv + a
    sp = 0
    for a in pres:
        sp = sp + a
    if prec == "ref32":
        f = np.float32
        sv = ((sv - f(vave)) / f(vstd)).astype(f)
        sp = ((sp - f(pave)) / f(pstd)).astype(f)
        mixed = (f(1 - alpha) * sv + f(alpha) * sp).astype(f)
    else:
        mixed = (1 - alpha) * (sv - vave) / vstd + alpha * (sp - pave) / pstd
    best = mixed.max()
    trans = tuple(int(v) for v in np.unravel_index(mixed.argmax(), mixed.shape))
    return best, trans


# --------------------------------------------------------------------------
# Per-rotation scoring and the search loop (d), refinement (e)
# --------------------------------------------------------------------------

def rot_and_search_fft(data, vec, angle, tspecs, target: MapGrid, mode="VecProduct",
                       prec="ref32", fma=True):
    """``rot_and_search_fft`` (VESPER_1_prob_new.py:1418-1459).  Returns the
    float64 score *before* the 1/N^3 scaling, the translation, and the
    rotated maps.  The CC/PCC rescale is applied here as in :1452-1457."""
    new_vec, new_data = rot_mrc(data, vec, angle, fma, prec)
    q = query_channels(new_data, new_vec, mode, target.ave)
    res = fft_correlate(tspecs, q, prec)
    best, trans = find_best_trans_list(res)
    score = np.float64(best)               # numpy-1.21: float32 scalar -> float64
    if mode == "CC":
        score = score * (1.0 / target.std ** 2)
    if mode == "PCC":
        score = score * (1.0 / target.std_norm_ave ** 2)
    return score, trans, new_vec, new_data


def refine_angles(angle):
    """+-5 degree 3x3x3 neighbourhood of one top-N row, filtered then wrapped
    (VESPER_1_prob_new.py:694-711)."""
    a = angle
    al = np.array(np.meshgrid([a[0] - 5, a[0], a[0] + 5],
                              [a[1] - 5, a[1], a[1] + 5],
                              [a[2] - 5, a[2], a[2] + 5])).T.reshape(-1, 3)
    al = al[(al[:, 0] < 360) & (al[:, 1] < 360) & (al[:, 2] < 180)]
    al[al < 0] += 360
    return al


def stable_topn(scores, topn):
    """``sorted(..., key=score, reverse=True)[:TopN]`` (VESPER_1_prob_new.py:664):
    Python's sort is stable and ``reverse=True`` keeps the original order of
    equal keys, so the order is (score descending, row ascending)."""
    order = sorted(range(len(scores)), key=lambda i: scores[i], reverse=True)
    return order[:topn]


def recheck_rows(scores, top, rel=1e-4):
    """Rows whose rank at the top-N cutoff is re-decided in FP64: every row
    scoring >= s_N - rel * max(|s_N|, |s_1|), s_1 and s_N the first and last
    scores of the top N, in row order.  (The FP64 re-check of the north star;
    the reference itself ranks by its float32 scores.)"""
    s1, sn = scores[top[0]], scores[top[-1]]
    thr = sn - rel * max(abs(sn), abs(s1))
    return [i for i in range(len(scores)) if scores[i] >= thr]


def search_map_fft(target: MapGrid, search: MapGrid, TopN=10, ang=30, mode="VecProduct",
                   prec="ref32", fma=True, rows=None, refine=True, fp64_recheck=False):
    """``search_map_fft`` (VESPER_1_prob_new.py:568-795) without printing.

    Returns dict(angles, scores (1/N^3-scaled float64), trans, topn (row ids),
    refined: list of dicts with angle/vec_score/vec_trans/vec/data).
    fp64_recheck: the rows near the top-N cutoff (recheck_rows) are re-scored
    with prec="ref64" and the top N taken by those scores (ties: row order);
    every refinement row is scored with "ref64"."""
    rd3 = 1.0 / (target.xdim ** 3)
    tspecs = target_spectra(target_channels(target, mode))
    angles = calc_angle_comb(ang) if rows is None else np.asarray(rows, np.float64)
    scores = np.zeros(len(angles))
    trans = np.zeros((len(angles), 3), np.int64)
    for i, a in enumerate(angles):
        s, t, _, _ = rot_and_search_fft(search.data, search.vec, a, tspecs, target, mode, prec, fma)
        scores[i] = s * rd3
        trans[i] = t
    top = stable_topn(list(scores), TopN)
    out = dict(angles=angles, scores=scores, trans=trans, topn=top, refined=None,
               ave=np.mean(scores), std=np.std(scores))
    top_score = {i: scores[i] for i in top}
    top_trans = {i: tuple(trans[i]) for i in top}
    if fp64_recheck:
        near = recheck_rows(scores, top)
        s64 = {}
        for i in near:
            s, t, _, _ = rot_and_search_fft(search.data, search.vec, angles[i], tspecs, target, mode, "ref64", fma)
            s64[i] = (s * rd3, tuple(t))
        top = sorted(near, key=lambda i: (-s64[i][0], i))[:TopN]
        top_score = {i: s64[i][0] for i in top}
        top_trans = {i: s64[i][1] for i in top}
        out.update(topn=top, recheck=near, recheck_scores=s64)
    out["topn_scores"] = [top_score[i] for i in top]
    out["topn_trans"] = [top_trans[i] for i in top]
    if not refine:
        return out
    if ang > 5.0:
        ref_rows = np.concatenate([refine_angles(angles[i]) for i in top], axis=0)
        refined = []
        rprec = "ref64" if fp64_recheck else prec
        for a in ref_rows:
            s, t, nv, nd = rot_and_search_fft(search.data, search.vec, a, tspecs, target, mode, rprec, fma)
            if fp64_recheck:
                # the returned maps are the reference's float32 rotation
                nv, nd = rot_mrc(search.data, search.vec, a, fma, "ref32")
            refined.append(dict(angle=tuple(a), vec_score=s * rd3, vec_trans=t, vec=nv, data=nd))
        order = stable_topn([r["vec_score"] for r in refined], TopN)
        out["refine_rows"] = ref_rows
        out["refined"] = [refined[i] for i in order]
    else:
        out["refined"] = [dict(angle=tuple(angles[i]), vec_score=top_score[i],
                               vec_trans=top_trans[i]) for i in top]
    return out


# --------------------------------------------------------------------------
# Prob mode
# --------------------------------------------------------------------------

def rot_and_search_fft_prob(data, vec, dp, angle, tspecs, alpha, stats=None,
                            prec="ref32", fma=True):
    """``rot_and_search_fft_prob`` (VESPER_1_prob_new.py:1462-1527).
    ``dp`` = (p1, p2, p3, p4); ``tspecs`` = [X, Y, Z, P1..P4];
    ``stats`` = (vstd, vave, pstd, pave) or None (pass 1)."""
    r = rot_mrc_prob(data, vec, *dp, angle, fma=fma, prec=prec)
    nv, nd, q1, q2, q3, q4 = r
    rv = fft_correlate(tspecs[:3], [nv[..., 0], nv[..., 1], nv[..., 2]], prec)
    rp = fft_correlate(tspecs[3:], [q1, q2, q3, q4], prec)
    vs, vt = find_best_trans_list(rv)
    ps, pt = find_best_trans_list(rp)
    ms, mt = None, None
    if stats is not None:
        vstd, vave, pstd, pave = stats
        ms, mt = find_best_trans_mixed(rv, rp, alpha, vstd, vave, pstd, pave, prec)
    return vs, vt, ps, pt, ms, mt, nv, nd


def search_map_fft_prob(tp, target: MapGrid, search: MapGrid, sp, ang, alpha=0.0, TopN=10,
                        vave=-10, vstd=-10, pave=-10, pstd=-10, prec="ref32", fma=True,
                        refine=True):
    """``search_map_fft_prob`` (VESPER_1_prob_new.py:798-1141) without printing.
    ``tp`` = target probability MapGrids (4), ``sp`` = search ones (4)."""
    rd3 = 1.0 / (target.xdim ** 3)
    chans = [target.vec[..., 0], target.vec[..., 1], target.vec[..., 2]] + [p.data for p in tp]
    tspecs = target_spectra(chans)
    angles = calc_angle_comb(ang)
    dp = tuple(p.data for p in sp)
    if not (vave >= 0 and vstd >= 0 and pstd >= 0 and pave >= 0):     # :877
        vsc, psc = [], []
        for a in angles:
            vs, _, ps, _, _, _, _, _ = rot_and_search_fft_prob(search.data, search.vec, dp, a, tspecs,
                                                               0.0, None, prec, fma)
            vsc.append(np.float64(vs) * rd3)
            psc.append(np.float64(ps) * rd3)
        vstd, vave = np.std(vsc), np.mean(vsc)
        pstd, pave = np.std(psc), np.mean(psc)
    rows = []
    for a in angles:
        vs, vt, ps, pt, ms, mt, _, _ = rot_and_search_fft_prob(
            search.data, search.vec, dp, a, tspecs, alpha, (vstd, vave, pstd, pave), prec, fma)
        rows.append(dict(angle=a, vec_score=np.float64(vs) * rd3, vec_trans=vt,
                         prob_score=np.float64(ps) * rd3, prob_trans=pt,
                         mixed_score=np.float64(ms) * rd3, mixed_trans=mt))
    top = stable_topn([r["mixed_score"] for r in rows], TopN)
    out = dict(rows=rows, topn=top, stats=(vstd, vave, pstd, pave), refined=None)
    if not refine:
        return out
    if ang > 5.0:
        ref_rows = np.concatenate([refine_angles(rows[i]["angle"]) for i in top], axis=0)
        refined = []
        for a in ref_rows:
            vs, vt, ps, pt, ms, mt, nv, nd = rot_and_search_fft_prob(
                search.data, search.vec, dp, a, tspecs, alpha, (vstd, vave, pstd, pave), prec, fma)
            refined.append(dict(angle=tuple(a), vec_score=np.float64(vs) * rd3, vec_trans=vt,
                                vec=nv, data=nd, prob_score=np.float64(ps) * rd3, prob_trans=pt,
                                mixed_score=np.float64(ms) * rd3, mixed_trans=mt))
        order = stable_topn([r["mixed_score"] for r in refined], TopN)
        out["refined"] = [refined[i] for i in order]
    else:
        out["refined"] = [rows[i] for i in top]
    return out


# --------------------------------------------------------------------------
# Final scoring
# --------------------------------------------------------------------------

def get_score(target: MapGrid, sdata, svec, trans, prec="ref32"):
    """``get_score`` (VESPER_1_prob_new.py:1302-1393), including its
    ``d2 = where(d2 <= 0, 0, d1)`` line (:1357): CC sums d1^2 over voxels where
    both maps are positive.  Returns the printed quantities and sco_arr."""
    tdata, tvec = target.data, target.vec
    ave1 = target.ave
    pos = sdata[sdata > 0]
    ave2 = np.mean(pos)
    std1 = target.std
    std2 = np.linalg.norm(pos)
    pstd1 = target.std_norm_ave
    pstd2 = np.linalg.norm(pos - ave2)
    dim = tdata.shape[0]
    t = np.array(trans)
    for i in range(3):
        if trans[i] > 0.5 * dim:
            t[i] -= dim
    tp = np.array(np.meshgrid(np.arange(dim), np.arange(dim), np.arange(dim))).T.reshape(-1, 3)
    spos = tp + t
    total = np.count_nonzero(tdata[tp[:, 0], tp[:, 1], tp[:, 2]])
    ok = ((spos >= 0) & (spos < dim)).all(axis=1)
    tp, spos = tp[ok], spos[ok]
    d1 = tdata[tp[:, 0], tp[:, 1], tp[:, 2]]
    d2 = sdata[spos[:, 0], spos[:, 1], spos[:, 2]]
    d1 = np.where(d1 <= 0, 0.0, d1)
    d2 = np.where(d2 <= 0, 0.0, d1)
    pd1 = np.where(d1 <= 0, 0.0, d1 - ave1)
    pd2 = np.where(d2 <= 0, 0.0, d2 - ave2)
    cc = np.sum(np.multiply(d1, d2))
    pcc = np.sum(np.multiply(pd1, pd2))
    tz = tdata[tp[:, 0], tp[:, 1], tp[:, 2]] == 0
    tnz = tdata[tp[:, 0], tp[:, 1], tp[:, 2]] > 0
    snz = sdata[spos[:, 0], spos[:, 1], spos[:, 2]] > 0
    total += np.count_nonzero(tz & snz)
    tv = tvec[tp[:, 0], tp[:, 1], tp[:, 2]]
    sv = svec[spos[:, 0], spos[:, 1], spos[:, 2]]
    sco_arr = np.zeros_like(sdata)
    if prec == "ref32":
        sco = np.einsum("ij,ij->i", tv, sv)
    else:
        sco = np.einsum("ij,ij->i", tv.astype(np.float64), sv.astype(np.float64))
        sco_arr = sco_arr.astype(np.float64)
    sco_arr[spos[:, 0], spos[:, 1], spos[:, 2]] = sco
    score = np.sum(sco_arr)
    nm = int(np.count_nonzero(tnz & snz))
    return dict(overlap=float(nm) / float(total), Nm=nm, total=int(total),
                cc=cc / (std1 * std2), pcc=pcc / (pstd1 * pstd2), score=score,
                cc_raw=cc, pcc_raw=pcc, sco_arr=sco_arr)


def euler2rot(euler):
    """VESPER_1_prob_new.py:1549-1551."""
    return Rotation.from_euler("xyz", euler, degrees=True)


def convert_trans(cen1, cen2, r, trans, xwidth2, dim):
    """VESPER_1_prob_new.py:1554-1566."""
    trans = np.array(trans)
    for i in range(3):
        if trans[i] > 0.5 * dim:
            trans[i] -= dim
    cen2 = r.apply(cen2)
    return cen1 - (cen2 + trans * xwidth2)


# --------------------------------------------------------------------------
# Brute-force helpers (used by the oracle's own pin tests)
# --------------------------------------------------------------------------

def brute_circular_corr(t, q):
    """sum_v t[v] q[(v + s) mod N] for every shift s, by explicit loops."""
    n = t.shape[0]
    out = np.zeros(t.shape, np.float64)
    tv = t.astype(np.float64)
    qv = q.astype(np.float64)
    for s in itertools.product(range(n), repeat=3):
        out[s] = np.sum(tv * np.roll(qv, shift=(-s[0], -s[1], -s[2]), axis=(0, 1, 2)))
    return out
