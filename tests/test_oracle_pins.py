"""Pins of the oracle (oracle/vesper_ref.py, oracle/fastvec_ref.c) against
closed forms, brute force and invariants.  The reference ships no tests or
golden vectors, so these checks are what fixes the oracle; each is chosen so
that a plausible slip (transposed matrix, wrong sign, off-by-one window, lost
fold, reversed tie order) fails it."""
import itertools

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import vesper_ref as ref
from oracle.build_c import fastvec_c


# ------------------------------------------------------------------ geometry

def _brute_fast(n):
    f = n
    cnt = {}
    for p in (2, 3, 5, 7, 11, 13):
        while f % p == 0:
            cnt[p] = cnt.get(p, 0) + 1
            f //= p
    return f == 1 and cnt.get(11, 0) + cnt.get(13, 0) <= 1


def test_next_fast_len_bruteforce_and_notebook_evidence():
    for n in range(1, 800):
        m = ref.next_fast_len(n)
        assert _brute_fast(m) and m >= n
        assert all(not _brute_fast(k) for k in range(n, m))
    # Notebooks/VESPER_prob_rot.ipynb:1431, :1439 printed Nvox= 39 and Nvox= 55
    assert ref.next_fast_len(39) == 39 and ref.next_fast_len(55) == 55
    assert ref.next_fast_len(38) == 39 and ref.next_fast_len(17) == 18
    assert ref.next_fast_len(143) == 144          # 11*13 is excluded
    assert ref.next_fast_len(173) == 175


def test_angle_rows_closed_form():
    for a, count in ((30, 1008), (10, 24624), (5, 191808), (30.0, 1008)):
        rows = ref.calc_angle_comb(a)
        nx = len(np.arange(0, 360, a))
        nz = len(np.arange(0, 180 + 1e-9, a))
        assert len(rows) == count == nx * nx * nz
    rows = ref.calc_angle_comb(30)
    # z slowest, x middle, y fastest
    assert rows[:3].tolist() == [[0, 0, 0], [0, 30, 0], [0, 60, 0]]
    assert rows[12].tolist() == [30, 0, 0]
    assert rows[144].tolist() == [0, 0, 30]
    assert rows[-1].tolist() == [330, 330, 180]


def test_refine_rows_closed_form():
    r = ref.refine_angles(np.array([0.0, 0.0, 0.0]))
    assert len(r) == 27
    # z = -5 survives the z < 180 filter, then wraps to 355 (Q15)
    assert set(r[:, 2].tolist()) == {355.0, 0.0, 5.0}
    assert set(r[:, 0].tolist()) == {355.0, 0.0, 5.0}
    r = ref.refine_angles(np.array([355.0, 10.0, 180.0]))
    # x + 5 = 360 filtered; z < 180 drops both 180 and 185
    assert len(r) == 2 * 3 * 1
    # meshgrid order: z slowest, x middle, y fastest
    assert r[:3].tolist() == [[350, 5, 175], [350, 10, 175], [350, 15, 175]]


def test_flip_matrix_is_zyx_euler_flipped():
    rng = np.random.default_rng(0)
    for _ in range(20):
        ang = rng.uniform(0, 360, 3)
        f = ref.flip_matrix_f32(ang).astype(np.float64)
        m = Rotation.from_euler("ZYX", ang, degrees=True).as_matrix()
        assert np.allclose(f, m[::-1, ::-1], atol=1e-7)
        # flip(M) = P M P = Rx(-a) Ry(-b) Rz(-c) is the inverse of the extrinsic
        # xyz rotation euler2rot uses for output: gathering the source at F n
        # moves the query content by R_xyz (VESPER_1_prob_new.py:1549-1566)
        r = Rotation.from_euler("xyz", ang, degrees=True).as_matrix()
        assert np.allclose(f, r.T, atol=1e-7)
        assert not np.allclose(f, r, atol=1e-3)
    z = ref.flip_matrix_f32([90, 0, 0])
    # a 90 degree turn has 6 zero entries; scipy leaves ~6e-17 residues that
    # the |m| <= 1e-15 rule zeroes (VESPER_1_prob_new.py:307)
    assert np.count_nonzero(z == 0) == 6
    raw = Rotation.from_euler("ZYX", [90, 0, 0], degrees=True).as_matrix()
    assert np.count_nonzero(raw == 0) < 6


def test_fma32_is_correctly_rounded():
    rng = np.random.default_rng(1)
    a = rng.standard_normal(20000).astype(np.float32)
    b = rng.standard_normal(20000).astype(np.float32)
    c = (rng.standard_normal(20000) * 1e-3).astype(np.float32)
    got = ref._fma32(a, b, c)
    from fractions import Fraction
    for i in range(0, 20000, 997):
        exact = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        lo = np.float32(float(exact))
        cands = [np.nextafter(lo, np.float32(-np.inf)), lo, np.nextafter(lo, np.float32(np.inf))]
        best = min(cands, key=lambda v: (abs(Fraction(float(v)) - exact), int(np.float32(v).view(np.int32)) & 1))
        assert got[i] == best


# ------------------------------------------------------------------ gather

def _rand_maps(rng, n, density=0.5):
    data = rng.uniform(0.1, 1.0, (n, n, n)) * (rng.uniform(size=(n, n, n)) < density)
    vec = rng.standard_normal((n, n, n, 3)).astype(np.float32)
    vec *= (data != 0)[..., None]
    return data, vec


@pytest.mark.parametrize("n", [8, 10])
def test_gather_identity_even(n):
    rng = np.random.default_rng(n)
    data, vec = _rand_maps(rng, n)
    nv, nd = ref.rot_mrc(data, vec, [0, 0, 0])
    assert np.array_equal(nd, data) and np.array_equal(nv, vec)


@pytest.mark.parametrize("n", [7, 9])
def test_gather_identity_odd_fold(n):
    """Odd N: plane h <- rows {h, h+1}, the later (h+1) wins if non-zero;
    plane d > h <- d+1; plane N-1 is never written (Q9)."""
    rng = np.random.default_rng(n)
    data, vec = _rand_maps(rng, n, density=0.7)
    nv, nd = ref.rot_mrc(data, vec, [0, 0, 0])
    h = (n - 1) // 2
    src = list(range(h)) + [None] + list(range(h + 2, n))
    exp = np.zeros_like(data)
    for x, y, z in itertools.product(range(n), repeat=3):
        cand = []
        for d in (x, y, z):
            if d < h:
                cand.append([d])
            elif d == h:
                cand.append([h + 1, h])
            elif d < n - 1:
                cand.append([d + 1])
            else:
                cand.append([])
        # last row in (z, x, y) order that has density
        for cz in cand[2]:
            done = False
            for cx in cand[0]:
                for cy in cand[1]:
                    if data[cx, cy, cz] != 0:
                        exp[x, y, z] = data[cx, cy, cz]
                        done = True
                        break
                if done:
                    break
            if done:
                break
    assert np.array_equal(nd, exp)
    assert not nd[n - 1].any() and not nd[:, n - 1].any() and not nd[:, :, n - 1].any()
    del src


def test_gather_90_degree_is_exact_permutation():
    """A 90 degree turn about one axis is a permutation with a known
    direction; a transposed rotation or a sign slip maps voxels elsewhere."""
    n = 8
    rng = np.random.default_rng(3)
    data, vec = _rand_maps(rng, n, density=1.0)
    for ang in ([90, 0, 0], [0, 90, 0], [0, 0, 90], [90, 90, 0]):
        nv, nd = ref.rot_mrc(data, vec, ang)
        f = ref.flip_matrix_f32(ang).astype(np.float64)
        c = n / 2
        # destination voxel i takes source floor(F (i - c) + c)
        for x, y, z in itertools.product(range(n), repeat=3):
            o = f @ (np.array([x, y, z]) - c) + c
            if np.all(o >= 0) and np.all(o < n):
                s = tuple(np.floor(o + 1e-9).astype(int))
                assert nd[x, y, z] == data[s]
                assert np.allclose(nv[x, y, z], vec[s] @ f, atol=1e-6)
                # the moved vector keeps its source vector's norm
                assert np.isclose(np.linalg.norm(nv[x, y, z]), np.linalg.norm(vec[s]), atol=1e-6)
    # a rotation of 90 about x: y -> z; check one voxel explicitly against the
    # convention flip(M) = R_xyz so a transposed F would fail
    nv, nd = ref.rot_mrc(data, vec, [90, 0, 0])
    f = ref.flip_matrix_f32([90, 0, 0])
    assert f[1, 2] == 1 and f[2, 1] == -1


def test_gather_equivalent_euler_triples_identical():
    """(a, b, c) and (a+180, 180-b, c+180) are the same rotation
    (VESPER_1.ipynb:1275-1276 shows bit-identical scores for such pairs)."""
    n = 10
    rng = np.random.default_rng(4)
    data, vec = _rand_maps(rng, n)
    a = ref.flip_matrix_f32([90, 30, 180]).astype(np.float64)
    b = ref.flip_matrix_f32([270, 150, 0]).astype(np.float64)
    assert np.allclose(a, b, atol=1e-7)
    # the gathered maps: same voxels, same data, vectors within float32
    # rounding of the two (not bit-equal) matrices; and a different rotation
    # gives a different map
    va, da = ref.rot_mrc(data, vec, [90, 30, 180])
    vb, db = ref.rot_mrc(data, vec, [270, 150, 0])
    assert np.array_equal(da != 0, db != 0) and (da != 0).sum() > 50
    agree = np.mean(da == db)
    assert agree > 0.999, agree
    same = (da == db) & (da != 0)
    assert np.allclose(va[same], vb[same], atol=1e-5)
    _, dc = ref.rot_mrc(data, vec, [90, 30, 170])
    assert np.mean(dc == da) < agree


def test_gather_norm_preserved_random():
    n = 12
    rng = np.random.default_rng(5)
    data, vec = _rand_maps(rng, n)
    nrm = vec / np.maximum(np.linalg.norm(vec, axis=-1, keepdims=True), 1e-30)
    nrm = nrm.astype(np.float32)
    nv, nd = ref.rot_mrc(data, nrm, rng.uniform(0, 360, 3))
    m = nd != 0
    assert np.allclose(np.linalg.norm(nv[m], axis=1), 1.0, atol=1e-5)


# ------------------------------------------------------------------ FFT correlation

@pytest.mark.parametrize("n", [5, 6])
def test_fft_correlate_matches_bruteforce(n):
    rng = np.random.default_rng(n)
    t = rng.standard_normal((n, n, n))
    q = rng.standard_normal((n, n, n)).astype(np.float32)
    spec = ref.target_spectra([t])
    brute = ref.brute_circular_corr(t, q) * n ** 3
    got64 = ref.fft_correlate(spec, [q], "ref64")[0]
    got32 = ref.fft_correlate(spec, [q], "ref32")[0]
    assert np.allclose(got64, brute, rtol=1e-10, atol=1e-9)
    assert np.allclose(got32, brute, rtol=1e-4, atol=1e-3 * np.abs(brute).max())


def test_translated_self_copy_argmax_is_the_shift():
    n = 12
    rng = np.random.default_rng(7)
    t = rng.standard_normal((n, n, n))
    shift = (3, 7, 10)
    q = np.roll(t, shift, axis=(0, 1, 2)).astype(np.float32)
    res = ref.fft_correlate(ref.target_spectra([t]), [q], "ref32")
    best, trans = ref.find_best_trans_list(res)
    assert trans == shift
    assert np.isclose(best / n ** 3, np.sum(t * t), rtol=1e-5)


def test_argmax_first_index_on_ties():
    a = np.ones((4, 4, 4), np.float32)
    best, trans = ref.find_best_trans_list([a, a])
    assert trans == (0, 0, 0) and best == 2
    a[1, 2, 3] = 5
    a[2, 0, 0] = 5
    assert ref.find_best_trans_list([a])[1] == (1, 2, 3)


def test_mixed_score_by_voxel_loop_and_hand_value():
    """find_best_trans_mixed (VESPER_1_prob_new.py:1530-1545): (1 - alpha)
    weights the vector z-score and alpha the probability one.  Checked by a
    voxel loop with a scan for the first maximum, and on a hand-computed
    voxel where swapping the weights or the statistics changes the value."""
    rng = np.random.default_rng(8)
    v = [rng.standard_normal((4, 4, 4)).astype(np.float32) for _ in range(3)]
    p = [rng.standard_normal((4, 4, 4)).astype(np.float32) for _ in range(4)]
    alpha, vstd, vave, pstd, pave = 0.3, 2.0, 0.5, 4.0, -1.0
    best, bi = -np.inf, None
    for idx in itertools.product(range(4), repeat=3):
        zv = (sum(float(x[idx]) for x in v) - vave) / vstd
        zp = (sum(float(x[idx]) for x in p) - pave) / pstd
        val = (1 - alpha) * zv + alpha * zp
        if val > best:                      # strict: the first maximum in C order wins
            best, bi = val, idx
    got, trans = ref.find_best_trans_mixed(v, p, alpha, vstd, vave, pstd, pave, "ref64")
    assert np.isclose(got, best, rtol=1e-12) and trans == bi
    b32, t32 = ref.find_best_trans_mixed(v, p, alpha, vstd, vave, pstd, pave, "ref32")
    assert abs(b32 - best) < 1e-5 and t32 == bi
    # hand value: vector sum 5 -> z = (5 - 1) / 2 = 2; probability sum 1 ->
    # z = (1 - 1) / 4 = 0; mixed = (1 - 0.25) * 2 + 0.25 * 0 = 1.5
    vv = [np.zeros((2, 2, 2), np.float32) for _ in range(3)]
    pp = [np.zeros((2, 2, 2), np.float32) for _ in range(4)]
    vv[0][1, 0, 1], vv[2][1, 0, 1] = 2.0, 3.0
    pp[3][1, 0, 1] = 1.0
    for prec in ("ref64", "ref32"):
        m, t = ref.find_best_trans_mixed(vv, pp, 0.25, 2.0, 1.0, 4.0, 1.0, prec)
        assert t == (1, 0, 1) and m == 1.5


# ------------------------------------------------------------------ top-N

def test_stable_topn_ties_keep_row_order():
    s = [1.0, 3.0, 3.0, 2.0, 3.0, 0.5]
    assert ref.stable_topn(s, 4) == [1, 2, 4, 3]


# ------------------------------------------------------------------ get_score

def _score_brute(tdata, tvec, t_ave, sdata, svec, trans):
    n = tdata.shape[0]
    t = [v - n if v > 0.5 * n else v for v in trans]
    pos = sdata[sdata > 0]
    ave2 = pos.mean()
    nm = total_extra = 0
    cc = pcc = dot = 0.0
    for x, y, z in itertools.product(range(n), repeat=3):
        sx, sy, sz = x + t[0], y + t[1], z + t[2]
        if not (0 <= sx < n and 0 <= sy < n and 0 <= sz < n):
            continue
        d1 = tdata[x, y, z]
        ds = sdata[sx, sy, sz]
        if d1 > 0 and ds > 0:
            nm += 1
            cc += d1 * d1                          # the d2 = d1 line (:1357)
            pcc += (d1 - t_ave) * (d1 - ave2)
        if d1 == 0 and ds > 0:
            total_extra += 1
        dot += float(np.dot(tvec[x, y, z].astype(np.float64), svec[sx, sy, sz].astype(np.float64)))
    total = np.count_nonzero(tdata) + total_extra
    return nm, total, cc, pcc, dot


@pytest.mark.parametrize("trans", [(0, 0, 0), (2, 5, 1), (7, 0, 6)])
def test_get_score_bruteforce(trans):
    n = 8
    rng = np.random.default_rng(sum(trans))
    td, tv = _rand_maps(rng, n)
    sd, sv = _rand_maps(rng, n)
    st = ref.map_stats(td)
    t = ref.MapGrid(data=td, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=tv,
                    ave=st["ave"], std=st["std"], std_norm_ave=st["std_norm_ave"])
    r = ref.get_score(t, sd, sv, trans, "ref64")
    nm, total, cc, pcc, dot = _score_brute(td, tv, st["ave"], sd, sv, trans)
    assert r["Nm"] == nm and r["total"] == total
    assert np.isclose(r["cc_raw"], cc) and np.isclose(r["pcc_raw"], pcc) and np.isclose(r["score"], dot)


def test_fft_score_equals_get_score_dot_without_wrap():
    """Notebooks/VESPER_1.ipynb:1273 vs :1328: the FFT score x 1/N^3 equals
    get_score's DOT when the overlap does not wrap."""
    n = 16
    rng = np.random.default_rng(11)
    tv = np.zeros((n, n, n, 3), np.float32)
    td = np.zeros((n, n, n))
    tv[4:10, 4:10, 4:10] = rng.standard_normal((6, 6, 6, 3))
    td[4:10, 4:10, 4:10] = 1.0
    sv = np.roll(tv, (2, 1, 3), axis=(0, 1, 2))
    sd = np.roll(td, (2, 1, 3), axis=(0, 1, 2))
    st = ref.map_stats(td)
    t = ref.MapGrid(data=td, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=tv,
                    ave=st["ave"], std=st["std"], std_norm_ave=st["std_norm_ave"], xdim=n)
    spec = ref.target_spectra(ref.target_channels(t, "VecProduct"))
    res = ref.fft_correlate(spec, [sv[..., 0], sv[..., 1], sv[..., 2]], "ref32")
    best, trans = ref.find_best_trans_list(res)
    assert trans == (2, 1, 3)
    r = ref.get_score(t, sd, sv, trans, "ref32")
    assert np.isclose(np.float64(best) / n ** 3, r["score"], rtol=1e-5)


# ------------------------------------------------------------------ fastVEC

def _small_src(rng, s=14):
    src = np.zeros((s, s, s), np.float32)
    m = rng.uniform(size=src.shape) < 0.15
    src[m] = rng.uniform(0.5, 1.5, m.sum()).astype(np.float32)
    return src


def test_fastvec_c_matches_numpy_direct(fastvec_ref):
    rng = np.random.default_rng(12)
    src = _small_src(rng)
    so = np.array([1.0, -2.0, 0.5], np.float32)
    g = ref.MapGrid(data=src, xwidth=np.float32(1.0), orig=so)
    n, w = 8, 2.0
    do = np.array([0.3, -1.7, 1.1])
    d_np, v_np = ref.fastvec(g, n, w, do, dreso=3.0)
    d_c, v_c = fastvec_c(fastvec_ref, src, 1.0, so, n, w, do, 3.0)
    assert np.array_equal(d_np != 0, d_c != 0) and (d_c != 0).sum() > 20
    assert np.allclose(d_np, d_c, rtol=1e-12, atol=0)
    assert np.max(np.abs(v_np - v_c)) <= 2e-7


def test_fastvec_single_voxel_points_at_it(fastvec_ref):
    """One source voxel: every computed destination vector is the unit vector
    towards it, and data = its Gaussian weight (closed form)."""
    s = 20
    src = np.zeros((s, s, s), np.float32)
    src[9, 11, 8] = 2.0
    so = np.zeros(3, np.float32)
    n, w, dreso = 10, 1.0, 4.0
    do = np.array([4.0, 6.0, 3.0])
    # only destination voxels sitting on non-zero source are computed: make the
    # source non-zero under a few destination voxels by a tiny background
    src2 = src.copy()
    src2[4:14, 6:16, 3:13] += np.float32(1e-30)
    d, v = fastvec_c(fastvec_ref, src2, 1.0, so, n, w, do, dreso)
    fsiv = 1.0 / ((dreso / 1.0) * 0.5) ** 2
    for (x, y, z) in [(0, 0, 0), (5, 5, 5), (9, 2, 7)]:
        pos = np.array([x, y, z]) * w + do
        r = np.array([9, 11, 8]) - pos
        if np.all(np.abs(r) <= 2 * dreso):
            assert np.isclose(d[x, y, z], 2.0 * np.exp(-1.5 * (r @ r) * fsiv), rtol=1e-6)
            if np.linalg.norm(r) > 0:
                assert np.allclose(v[x, y, z], r / np.linalg.norm(r), atol=1e-5)


def test_fastvec_window_edges(fastvec_ref):
    """The window runs [trunc(pos - fmaxd), trunc(pos + fmaxd + 1)): a source
    voxel exactly fmaxd away is included, one further is not."""
    s = 30
    src = np.zeros((s, s, s), np.float32)
    src[10, 10, 10] = 1.0
    dreso = 4.0                                 # fmaxd = 8
    for off, inside in ((8, True), (9, False)):
        src2 = src.copy()
        src2[10 + off, 10, 10] = 1.0
        d, _ = fastvec_c(__import__("oracle.build_c", fromlist=["x"]).load(), src2, 1.0, np.zeros(3, np.float32),
                         1, 1.0, np.array([10.0, 10.0, 10.0]), dreso)
        fsiv = 1.0 / 4.0
        exp = 1.0 + (np.exp(-1.5 * off * off * fsiv) if inside else 0.0)
        assert np.isclose(d[0, 0, 0], exp, rtol=1e-12)


# ------------------------------------------------------------------ map preparation

def _laplacian_loops(a):
    """laplacian_filter as the reference's explicit loops (VESPER_1_prob_new.py:
    509-525): only voxels > 0, neighbours outside the grid contribute nothing."""
    nx, ny, nz = a.shape
    out = np.zeros_like(a)
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                if a[x][y][z] > 0:
                    v = -6.0 * a[x][y][z]
                    if x + 1 < nx:
                        v += a[x + 1][y][z]
                    if x - 1 >= 0:
                        v += a[x - 1][y][z]
                    if y + 1 < ny:
                        v += a[x][y + 1][z]
                    if y - 1 >= 0:
                        v += a[x][y - 1][z]
                    if z + 1 < nz:
                        v += a[x][y][z + 1]
                    if z - 1 >= 0:
                        v += a[x][y][z - 1]
                    out[x][y][z] = v
    return out


def test_laplacian_filter_loops_and_closed_form():
    rng = np.random.default_rng(30)
    a = rng.standard_normal((7, 6, 5))            # negatives, boundary voxels, positives
    a[rng.uniform(size=a.shape) < 0.2] = 0.0
    assert np.array_equal(ref.laplacian_filter(a), _laplacian_loops(a))
    # one positive voxel: -6 v there; its zero neighbours stay 0 (only
    # positive voxels are filtered); a corner voxel sees no wrapped neighbour
    b = np.zeros((5, 5, 5))
    b[2, 2, 2] = 1.5
    b[0, 0, 0] = 2.0
    b[4, 4, 4] = -3.0
    got = ref.laplacian_filter(b)
    exp = np.zeros_like(b)
    exp[2, 2, 2] = -9.0
    exp[0, 0, 0] = -12.0
    assert np.array_equal(got, exp)
    # two positive neighbours see each other
    c = np.zeros((4, 4, 4))
    c[1, 2, 3], c[2, 2, 3] = 1.0, 4.0
    got = ref.laplacian_filter(c)
    assert got[1, 2, 3] == -6.0 + 4.0 and got[2, 2, 3] == -24.0 + 1.0 and np.count_nonzero(got) == 2


def test_set_vox_size_closed_form():
    """mrc_set_vox_size (VESPER_1_prob_new.py:59-118) by hand.  Grid 20 x 30
    x 40 at 1.5 A, origin (1, -2, 0.5), centre (10, 15, 20) voxels.  The
    farthest voxel above the threshold sits at (8, 18, 26): offset (-2, 3, 6),
    dmax = 7 voxels = 10.5 A; at 2 A per voxel int(2 * 10.5 / 2) = 10 -> N =
    10; new centre = centre * 1.5 + origin = (16, 20.5, 30.5); new origin =
    new centre - 0.5 * 10 * 2 = (6, 10.5, 20.5)."""
    d = np.zeros((20, 30, 40), np.float32)
    d[8, 18, 26] = 0.9
    d[10, 15, 20] = 0.5
    d[19, 29, 39] = 0.2             # farther, but below the threshold
    d[0, 0, 0] = -1.0               # negative: below the threshold
    m = ref.MapGrid(data=d, xwidth=np.float32(1.5), orig=np.array([1.0, -2.0, 0.5], np.float32))
    src, dst = ref.set_vox_size(m, 0.3, 2.0)
    assert dst.xdim == 10
    assert dst.extra["dmax_vox"] == 7.0 and dst.extra["dmax_printed"] == str(7.0 / 1.5)
    assert np.array_equal(dst.cent, [16.0, 20.5, 30.5])
    assert np.array_equal(dst.orig, [6.0, 10.5, 20.5])
    assert np.count_nonzero(src.data) == 2 and src.data[19, 29, 39] == 0 and src.data[0, 0, 0] == 0
    # a negative threshold leaves the data alone and thresholds at 0 (:73-79):
    # the 0.2 voxel now counts, offset (9, 14, 19) -> dmax = sqrt(638) voxels
    src, dst = ref.set_vox_size(m, -0.1, 2.0)
    assert src.data[19, 29, 39] == np.float32(0.2) and src.data[0, 0, 0] == 0
    assert dst.xdim == ref.next_fast_len(int(2 * np.sqrt(638.0) * 1.5 / 2.0))
    # int() truncates: dmax 7 at 1.5 A and 1.9 A per voxel gives 11.05 -> 11
    m2 = ref.MapGrid(data=d, xwidth=np.float32(1.5), orig=np.zeros(3, np.float32))
    assert ref.set_vox_size(m2, 0.3, 1.9)[1].xdim == 11
    assert ref.set_vox_size(m2, 0.3, 1.3)[1].xdim == 16        # int(16.15) = 16


def test_unify_dims_closed_form():
    """main.py:166-178: the smaller grid takes the larger N and its origin
    becomes centre - 0.5 * s * N."""
    s = 2.0
    a = ref.MapGrid(data=np.zeros((2, 2, 2)), xwidth=s, orig=np.array([0.0, 0.0, 0.0]),
                    cent=np.array([30.0, 40.0, 50.0]), xdim=16)
    b = ref.MapGrid(data=np.zeros((2, 2, 2)), xwidth=s, orig=np.array([9.0, 9.0, 9.0]),
                    cent=np.array([1.0, 2.0, 3.0]), xdim=21)
    assert ref.unify_dims(a, b, s) == 21 and a.xdim == 21 and b.xdim == 21
    assert np.array_equal(a.orig, [30.0 - 21.0, 40.0 - 21.0, 50.0 - 21.0])
    assert np.array_equal(b.orig, [9.0, 9.0, 9.0])


def test_map_stats_norms_not_std():
    """VESPER_1_prob_new.py:157-161: `std` is the L2 norm of the positive
    voxels (not a standard deviation), std_norm_ave the L2 norm of their
    deviations from their mean; the sum and count include negatives."""
    d = np.array([[[3.0, 4.0], [0.0, -1.0]]])
    st = ref.map_stats(d)
    assert st["dsum"] == 6.0 and st["Nact"] == 3
    assert st["ave"] == 3.5
    assert st["std"] == 5.0                         # sqrt(9 + 16); a std would be 0.5
    assert np.isclose(st["std_norm_ave"], np.sqrt(0.5), rtol=1e-15)


def test_mode_channels_closed_form():
    """O/C/P/L channels (VESPER_1_prob_new.py:546-558 query, :608-615
    target).  PCC subtracts the given average -- the target's, also for the
    query (:1450)."""
    d = np.array([-1.0, 0.0, 0.5, 2.0]).reshape(1, 1, 4)
    vec = np.arange(12, dtype=np.float32).reshape(1, 1, 4, 3)
    t = ref.MapGrid(data=d, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=vec, ave=0.25)
    exp = {"Overlap": [0.0, 0.0, 1.0, 1.0], "CC": [0.0, 0.0, 0.5, 2.0], "PCC": [0.0, 0.0, 0.25, 1.75],
           "Laplacian": [0.0, 0.0, -3.0 + 2.0, -12.0 + 0.5]}
    for mode, want in exp.items():
        (tc,) = ref.target_channels(t, mode)
        (qc,) = ref.query_channels(d, vec, mode, 0.25)
        assert np.array_equal(tc.ravel(), want) and np.array_equal(qc.ravel(), want), mode
    (qp,) = ref.query_channels(d, vec, "PCC", 1.0)        # the average handed in is the one used
    assert np.array_equal(qp.ravel(), [0.0, 0.0, -0.5, 1.0])
    tv = ref.target_channels(t, "VecProduct")
    assert all(np.array_equal(tv[k], vec[..., k]) for k in range(3))


def _fold_expected(p, n):
    """Odd-N identity gather of one channel with its own mask: plane h takes
    row h + 1 if that voxel is non-zero, else row h (the later row wins only
    when it passes its filter); planes above h take d + 1; plane N-1 stays 0."""
    h = (n - 1) // 2
    exp = np.zeros_like(p)

    def cands(d):
        if d < h:
            return [d]
        if d == h:
            return [h + 1, h]
        if d < n - 1:
            return [d + 1]
        return []
    for x, y, z in itertools.product(range(n), repeat=3):
        for cz in cands(z):
            hit = False
            for cx in cands(x):
                for cy in cands(y):
                    if p[cx, cy, cz] != 0:
                        exp[x, y, z] = p[cx, cy, cz]
                        hit = True
                        break
                if hit:
                    break
            if hit:
                break
    return exp


def test_rot_mrc_prob_own_masks():
    """rot_mrc_prob (VESPER_1_prob_new.py:1245-1297) gathers each
    probability channel with its own non-zero mask, not the density's: a
    probability voxel where the density is 0 still moves, and at the odd-N
    centre plane each channel picks its own surviving row."""
    rng = np.random.default_rng(31)
    for n in (8, 9):
        data, vec = _rand_maps(rng, n, density=0.4)
        probs = [(rng.uniform(0.1, 1, (n, n, n)) * (rng.uniform(size=(n, n, n)) < 0.5)).astype(np.float32)
                 for _ in range(4)]
        # probability where the density is zero
        assert any(np.any((p != 0) & (data == 0)) for p in probs)
        out = ref.rot_mrc_prob(data, vec, *probs, [0, 0, 0])
        nv, nd = ref.rot_mrc(data, vec, [0, 0, 0])
        assert np.array_equal(out[0], nv) and np.array_equal(out[1], nd)
        for k in range(4):
            exp = probs[k] if n % 2 == 0 else _fold_expected(probs[k], n)
            assert np.array_equal(out[2 + k], exp), (n, k)
    # 90 degrees about x on an even grid: each channel is the same
    # permutation of itself, its zeros and the density's notwithstanding
    n = 8
    data, vec = _rand_maps(rng, n, density=0.3)
    probs = [(rng.uniform(0.1, 1, (n, n, n)) * (rng.uniform(size=(n, n, n)) < 0.5)).astype(np.float32)
             for _ in range(4)]
    out = ref.rot_mrc_prob(data, vec, *probs, [90, 0, 0])
    f = ref.flip_matrix_f32([90, 0, 0]).astype(np.float64)
    moved = 0
    for x, y, z in itertools.product(range(n), repeat=3):
        o = f @ (np.array([x, y, z]) - n / 2) + n / 2
        if np.all(o >= 0) and np.all(o < n):
            sidx = tuple(np.floor(o + 1e-9).astype(int))
            for k in range(4):
                assert out[2 + k][x, y, z] == probs[k][sidx]
                moved += int(probs[k][sidx] != 0 and data[sidx] == 0)
        else:
            assert all(out[2 + k][x, y, z] == 0 for k in range(4))
    assert moved > 10


def test_fp64_recheck_equals_fp64_everywhere():
    """The FP64 re-check scores only the rows near the float32 top-N cutoff
    in FP64; its top N must equal the top N of scoring every row in FP64."""
    n = 10
    rng = np.random.default_rng(40)
    td, tv = _rand_maps(rng, n, density=0.4)
    sd, sv = _rand_maps(rng, n, density=0.4)
    st = ref.map_stats(td)
    t = ref.MapGrid(data=td, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=tv, ave=st["ave"],
                    std=st["std"], std_norm_ave=st["std_norm_ave"], xdim=n)
    s = ref.MapGrid(data=sd, xwidth=np.float32(1), orig=np.zeros(3, np.float32), vec=sv, xdim=n)
    rows = ref.calc_angle_comb(30)[::3]
    got = ref.search_map_fft(t, s, TopN=10, mode="VecProduct", rows=rows, refine=False, fp64_recheck=True)
    full = ref.search_map_fft(t, s, TopN=10, mode="VecProduct", rows=rows, refine=False, prec="ref64")
    assert got["topn"] == full["topn"]
    assert np.allclose(got["topn_scores"], [full["scores"][i] for i in full["topn"]], rtol=1e-12)
    assert got["topn_trans"] == [tuple(full["trans"][i]) for i in full["topn"]]
    assert set(full["topn"]) <= set(got["recheck"]) and len(got["recheck"]) < len(rows)
