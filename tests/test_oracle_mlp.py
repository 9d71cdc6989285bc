"""Pins of the proposal network's oracle (oracle/mlp_ref.py): analytic
gradients against central finite differences of the exact float64 network,
bf16 rounding against a direct bit-level rule, moments against closed forms."""
import numpy as np

from oracle import mlp_ref as mr


def _net(rng, inp=12, h=8):
    W = [rng.standard_normal((inp, h)) * 0.5, rng.standard_normal((h, h)) * 0.5, rng.standard_normal((h, h)) * 0.5]
    b = [rng.standard_normal(h) * 0.1 for _ in range(3)]
    return W, b, rng.standard_normal(h) * 0.5, np.array([0.1])


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((20, 12))
    y = rng.standard_normal(20)
    W, b, w4, b4 = _net(rng)
    g, loss = mr.grads(X, y, W, b, w4, b4, round_bf16=False)
    params = W + b + [w4, b4]

    def L():
        p, _ = mr.forward(X, W, b, w4, b4, round_bf16=False)
        return np.mean((p - y) ** 2)

    assert np.isclose(loss, L())
    eps = 1e-6
    for pi, (p, gp) in enumerate(zip(params, g)):
        flat = p.reshape(-1)
        for j in range(0, flat.size, max(1, flat.size // 7)):
            old = flat[j]
            flat[j] = old + eps
            lp = L()
            flat[j] = old - eps
            lm = L()
            flat[j] = old
            assert abs((lp - lm) / (2 * eps) - gp.reshape(-1)[j]) < 1e-6 * max(1.0, abs(gp.reshape(-1)[j])), (pi, j)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159265, 1e-30], np.float32)
    got = mr.bf16(vals)
    for v, g in zip(vals, got):
        u = np.float32(v).view(np.uint32)
        lo = np.uint32(u & 0xFFFF0000).view(np.float32)
        hi = np.uint32((u & 0xFFFF0000) + 0x10000).view(np.float32)
        cands = sorted([lo, hi], key=lambda c: (abs(float(c) - float(v)), int(np.float32(c).view(np.uint32) >> 16) & 1))
        assert g == cands[0]


def test_moments_closed_form():
    n = 8
    data = np.zeros((n, n, n))
    vec = np.zeros((n, n, n, 3), np.float32)
    data[6, 4, 4] = 2.0
    vec[6, 4, 4] = [0, 0, 1]
    m = mr.moments(data, vec)
    assert np.allclose(m[0:3], [2 / n, 0, 0])             # centroid (6 - 4) / n
    assert np.isclose(m[3], (2 / n) ** 2) and np.allclose(m[4:9], 0)
    assert np.allclose(m[9:12], [0, 0, 1])
    assert np.isclose(m[12 + 2], 2 / n)                    # r_x v_z
    assert np.isclose(m[23], 1.0) and np.isclose(m[31], 1 / n ** 3)
