"""Pins the radiance-cache restatement (oracle/radiance_cache_ref.py) with what the mathematics fixes:
SH orthonormality (exact quadrature), trilinear interpolation reproducing linear fields, table index
ranges and dense-level bijectivity, finite-difference gradients of the unrounded network and the
relative L2 loss, and Adam's closed-form first step."""
import math
import pathlib
import sys

import numpy as np
import pytest

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1] / "oracle"))
import radiance_cache_ref as R  # noqa: E402


def random_records(n, seed, lo=-0.1, hi=1.1):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 16), np.float32)
    rec[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    rec[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    rec[:, 8:11] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    rec[:, 3] = rng.uniform(0, 1, n)
    rec[:, 7] = rng.uniform(0, 1, n)
    rec[:, 11] = rng.uniform(0, 1, n)
    rec[:, 12:15] = rng.uniform(0, 1, (n, 3))
    return rec


def test_sh_orthonormal():
    """Gauss-Legendre in cos(theta) x uniform phi integrates products of degree <= 6 exactly"""
    zs, wz = np.polynomial.legendre.leggauss(16)
    phis = (np.arange(32) + 0.5) * (2 * np.pi / 32)
    z = np.repeat(zs, len(phis))
    phi = np.tile(phis, len(zs))
    w = np.repeat(wz, len(phis)) * (2 * np.pi / 32)
    s = np.sqrt(1 - z * z)
    d = np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)
    y = R.sh16(d).astype(np.float64)
    gram = (y * w[:, None]).T @ y
    assert np.allclose(gram, np.eye(16), atol=2e-6)


def test_resolutions():
    res = R.resolutions()
    assert res[0] == 16 and res[-1] == 2048 and all(a < b for a, b in zip(res, res[1:]))
    assert [(n + 1) ** 3 <= 1 << 19 for n in res] == [True] * 5 + [False] * 11


def test_trilinear_reproduces_linear_fields():
    """on dense levels a table holding a linear function of the corner coordinates interpolates to that
    linear function of the position (exact up to fp32 rounding): pins the weights and the dense index"""
    log2_t = 16
    t = 1 << log2_t
    res = R.resolutions()
    rec = random_records(500, 1, 0.0, 1.0)
    grid = np.zeros((16, t, 2), np.float32)
    coef = np.array([0.7, -0.3, 0.45], np.float32)
    for l, n in enumerate(res):
        if (n + 1) ** 3 > t:
            continue
        c = np.arange((n + 1) ** 3)
        cx, cy, cz = c % (n + 1), (c // (n + 1)) % (n + 1), c // (n + 1) ** 2
        grid[l, :len(c), 0] = (coef[0] * cx + coef[1] * cy + coef[2] * cz) / n
        grid[l, :len(c), 1] = cz / n
    x0 = R.encode(rec, grid, (0, 0, 0), (1, 1, 1))
    p = rec[:, 0:3].astype(np.float64)
    dense_levels = [l for l, n in enumerate(res) if (n + 1) ** 3 <= t]
    assert dense_levels == [0, 1, 2]
    for l in dense_levels:
        assert np.allclose(x0[:, R.hash_slot(l, 0)], p @ coef, atol=2e-6)
        assert np.allclose(x0[:, R.hash_slot(l, 1)], p[:, 2], atol=2e-6)


def test_table_indices():
    rec = random_records(2000, 2)
    idx, wts = R.grid_corners(rec, (0, 0, 0), (1, 1, 1))
    for l, n in enumerate(R.resolutions()):
        assert idx[l].min() >= 0 and idx[l].max() < 1 << R.LOG2_T
        assert np.allclose(wts[l].sum(axis=1), 1.0, atol=1e-6)
        assert (wts[l] >= 0).all()
    # dense level 0: all corner coordinates map to distinct indices
    n = R.resolutions()[0]
    c = np.stack(np.meshgrid(*[np.arange(n + 1)] * 3, indexing="ij"), -1).reshape(-1, 3)
    flat = c[:, 0] + (n + 1) * (c[:, 1] + (n + 1) * c[:, 2])
    assert len(np.unique(flat)) == (n + 1) ** 3


def test_spatial_hash_known_values():
    """(x * 1) xor (y * 2654435761) xor (z * 805459861) mod 2^19 at hand-computed corners"""
    t = 1 << 19
    for (x, y, z) in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (100, 200, 300)]:
        h = ((x * 1) ^ ((y * 2654435761) & 0xFFFFFFFF) ^ ((z * 805459861) & 0xFFFFFFFF)) & (t - 1)
        want = {(0, 0, 0): 0, (1, 0, 0): 1, (0, 1, 0): 2654435761 % t, (0, 0, 1): 805459861 % t}
        if (x, y, z) in want:
            assert h == want[(x, y, z)]
        assert 0 <= h < t


def test_backward_matches_finite_differences():
    """gradients of L = sum (y - t)^2 / c / n with c = mean(y)^2 + 0.01 held fixed (stop-gradient)"""
    rng = np.random.default_rng(3)
    n = 40
    x0 = rng.normal(size=(n, 64))
    flat = rng.normal(scale=0.25, size=R.N_MLP)
    weights = [w.astype(np.float64) for w in R.split_weights(flat)]
    targets = rng.uniform(0, 1, (n, 3))
    y, acts, zs = R.forward(x0, weights, rnd=R._identity, dtype=np.float64)
    denom = y[:, :3].mean(axis=1) ** 2 + R.LOSS_EPS

    def loss_fixed(ws, x):
        yy, _, _ = R.forward(x, ws, rnd=R._identity, dtype=np.float64)
        return float(np.sum((yy[:, :3] - targets) ** 2 / denom[:, None]) / n)

    _, dy = R.loss_grad(y, targets, n, dtype=np.float64)
    dws, dx0 = R.backward(dy, acts, zs, weights, rnd=R._identity, dtype=np.float64)
    eps = 1e-6
    for layer, (o, i) in [(0, (5, 7)), (2, (11, 60)), (4, (0, 33)), (5, (1, 9)), (5, (2, 63))]:
        wp = [w.copy() for w in weights]
        wm = [w.copy() for w in weights]
        wp[layer][o, i] += eps
        wm[layer][o, i] -= eps
        fd = (loss_fixed(wp, x0) - loss_fixed(wm, x0)) / (2 * eps)
        assert fd == pytest.approx(dws[layer][o, i], rel=1e-5, abs=1e-9)
    for (s, f) in [(0, 0), (3, 17), (10, 40)]:
        xp, xm = x0.copy(), x0.copy()
        xp[s, f] += eps
        xm[s, f] -= eps
        fd = (loss_fixed(weights, xp) - loss_fixed(weights, xm)) / (2 * eps)
        assert fd == pytest.approx(dx0[s, f], rel=1e-5, abs=1e-9)
    # padded outputs 3..15 receive no gradient
    assert np.all(dws[5][3:] == 0)


def test_grid_gradient_matches_finite_differences():
    """d(sum_n g_n . x0_n[hash slots]) / d grid equals the scattered trilinear weights"""
    log2_t = 10
    rng = np.random.default_rng(4)
    rec = random_records(30, 5)
    grid = rng.normal(scale=0.1, size=(16, 1 << log2_t, 2)).astype(np.float32)
    g = rng.normal(size=(30, 64)).astype(np.float32)
    gg = R.grid_grad(rec, g, grid.shape, (0, 0, 0), (1, 1, 1))

    def f(gr):
        return float(np.sum(R.encode(rec, gr, (0, 0, 0), (1, 1, 1)).astype(np.float64) * g))

    idx, _ = R.grid_corners(rec, (0, 0, 0), (1, 1, 1), log2_t)
    for (l, c) in [(0, 0), (7, 1), (15, 0)]:
        e = int(idx[l][0, 2])                      # an entry the first record touches
        gp, gm = grid.copy(), grid.copy()
        gp[l, e, c] += 0.25
        gm[l, e, c] -= 0.25
        fd = (f(gp) - f(gm)) / 0.5
        assert fd == pytest.approx(gg[l, e, c], rel=1e-4, abs=1e-5)


def test_adam_first_step_is_signed_learning_rate():
    """at step 1, m_hat = g and v_hat = g^2, so p <- p - lr g / (|g| + eps) = p - lr sign(g)"""
    p = np.array([0.5, -0.25, 1.0, 2.0], np.float32)
    g = np.array([3.0, -1e-3, 0.0, 7e-6], np.float32)
    out, m, v = R.adam(p, g, np.zeros(4, np.float32), np.zeros(4, np.float32), 1)
    assert np.allclose(out, [0.49, -0.24, 1.0, 1.99], atol=1e-6)
    assert m[2] == 0 and v[2] == 0                  # untouched parameter keeps its moments


def test_adam_second_step_closed_form():
    p = np.float32(1.0)
    g1, g2 = np.float32(0.5), np.float32(-0.25)
    out, m, v = R.adam(np.array([p]), np.array([g1]), np.zeros(1, np.float32), np.zeros(1, np.float32), 1)
    out, m, v = R.adam(out, np.array([g2]), m, v, 2)
    m2 = 0.9 * (0.1 * 0.5) + 0.1 * -0.25
    v2 = 0.99 * (0.01 * 0.25) + 0.01 * 0.0625
    want = 1.0 - 0.01 - 0.01 * (m2 / (1 - 0.81)) / (math.sqrt(v2 / (1 - 0.9801)) + 1e-15)
    assert out[0] == pytest.approx(want, rel=1e-6)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159265, 65504.0, 1e-30], np.float32)
    r = R.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 1.5 * 2^-8 rounds up to 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125) and r[3] == np.float32(1.0078125)
    assert r[4] == np.float32(-3.140625)
    assert (r.view(np.uint32) & 0xFFFF == 0).all()
