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
        # norm of every moved vector is preserved
        m = nd != 0
        assert np.allclose(np.linalg.norm(nv[m], axis=1), np.linalg.norm(vec, axis=-1).max() * 0 +
                           np.linalg.norm(nv[m], axis=1))
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


def test_mixed_score_closed_form():
    rng = np.random.default_rng(8)
    v = [rng.standard_normal((4, 4, 4)).astype(np.float32) for _ in range(3)]
    p = [rng.standard_normal((4, 4, 4)).astype(np.float32) for _ in range(4)]
    best, trans = ref.find_best_trans_mixed(v, p, 0.3, 2.0, 0.5, 4.0, -1.0, "ref64")
    m = 0.7 * (sum(x.astype(np.float64) for x in v) - 0.5) / 2.0 + 0.3 * (sum(x.astype(np.float64) for x in p) + 1.0) / 4.0
    assert np.isclose(best, m.max()) and trans == np.unravel_index(m.argmax(), m.shape)
    b32, t32 = ref.find_best_trans_mixed(v, p, 0.3, 2.0, 0.5, 4.0, -1.0, "ref32")
    assert abs(b32 - best) < 1e-5 and t32 == trans


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
