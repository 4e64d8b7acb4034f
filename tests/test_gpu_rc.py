"""GPU radiance cache against the numpy restatement (oracle/radiance_cache_ref.py): predictions, MLP and
hash-grid gradients, Adam's first step, and online fitting. The matrix cores accumulate fp32 in their own
order and the bf16 rounding of one activation can flip, so values are compared within tolerances stated
per test; structural outputs (which entries are touched, signs of clear gradients) compare exactly."""
import pathlib
import sys

import numpy as np
import pytest

import vierkant_amd as vk

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1] / "oracle"))
import radiance_cache_ref as R  # noqa: E402
from test_rc_oracle import random_records  # noqa: E402

pytestmark = pytest.mark.gpu

LOG2_T = 14                     # small table for the parity tests: levels 0..1 dense, 2..15 hashed
BMIN, BMAX = np.array([-1.0, 0.0, -2.0], np.float32), np.array([3.0, 2.0, 2.0], np.float32)


def inv_extent():
    return (np.float32(1.0) / (BMAX - BMIN)).astype(np.float32)


def world_records(n, seed):
    rec = random_records(n, seed)
    rec[:, 0:3] = BMIN + rec[:, 0:3] * (BMAX - BMIN)       # includes positions outside the box (clamped)
    return rec


@pytest.fixture()
def cache(ctx):
    c = vk.RadianceCache(ctx, vk.default_rc_params(log2_table=LOG2_T), seed=1)
    c.set_bounds(BMIN, BMAX)
    rng = np.random.default_rng(9)
    mlp = rng.normal(scale=0.2, size=c.n_mlp).astype(np.float32)
    grid = rng.uniform(-0.5, 0.5, size=(16, 1 << LOG2_T, 2)).astype(np.float32)
    c.set_params(mlp, grid)
    yield c
    c.close()


def stored_grid(grid):
    return grid.astype(np.float16).astype(np.float32)


def test_default_init(ctx):
    c = vk.RadianceCache(ctx, seed=5)
    mlp, grid = c.get_params()
    assert c.n_mlp == R.N_MLP and grid.shape == (16, 1 << 19, 2)
    assert np.abs(grid).max() <= 1e-4 and np.abs(grid).max() > 0.9e-4
    w6 = R.split_weights(mlp)[5]
    assert np.all(w6[3:] == 0) and np.all(w6[:3] != 0)
    c.close()


def test_query_matches_oracle(cache):
    rec = world_records(3001, 11)                           # not a multiple of the 256-sample block
    got = cache.query(rec)
    mlp, grid = cache.get_params()
    want = R.query(rec, stored_grid(grid), mlp, BMIN, inv_extent())
    scale = float(np.abs(want).max())
    err = np.abs(got - want)
    assert np.quantile(err, 0.99) <= 2e-3 * scale
    assert err.max() <= 3e-2 * scale
    assert np.array_equal(got == 0, want == 0) or np.mean((got == 0) != (want == 0)) < 2e-3


def test_gradients_match_oracle(cache):
    rec = world_records(1000, 12)
    tg = np.random.default_rng(13).uniform(0, 2, (1000, 3)).astype(np.float32)
    loss = cache.train(rec, tg, grad_only=True)
    assert cache.step() == 0
    g_mlp, g_grid = cache.get_gradients()
    mlp, grid = cache.get_params()
    want_loss, w_mlp, w_grid = R.train_gradients(rec, tg, stored_grid(grid), mlp, BMIN, inv_extent())
    assert loss == pytest.approx(want_loss, rel=1e-3)
    off = 0
    for (o, i) in R.LAYER_SHAPES:
        a, b = g_mlp[off: off + o * i], w_mlp[off: off + o * i]
        s = float(np.abs(b).max())
        assert np.abs(a - b).max() <= 2e-2 * s, f"layer at {off}"
        off += o * i
    assert np.all(R.split_weights(g_mlp)[5][3:] == 0)
    # grid gradients accumulate in fp16 (loss-scaled): only entries the oracle touches are non-zero, every
    # entry with a gradient above fp16 resolution is touched, and values agree to fp16 accumulation error
    s = float(np.abs(w_grid).max())
    assert not np.any((g_grid != 0) & (w_grid == 0))
    assert np.all(g_grid[np.abs(w_grid) > 1e-3 * s] != 0)
    assert np.abs(g_grid - w_grid).max() <= 2e-2 * s


def test_first_adam_step(cache):
    """step 1 moves every parameter with a clear gradient by exactly -lr sign(g)"""
    rec = world_records(700, 14)
    tg = np.random.default_rng(15).uniform(0, 2, (700, 3)).astype(np.float32)
    mlp0, grid0 = cache.get_params()
    _, w_mlp, w_grid = R.train_gradients(rec, tg, stored_grid(grid0), mlp0, BMIN, inv_extent())
    cache.train(rec, tg)
    assert cache.step() == 1
    mlp1, grid1 = cache.get_params()
    clear = np.abs(w_mlp) > 0.05 * np.abs(w_mlp).max()
    assert clear.sum() > 1000
    assert np.allclose((mlp1 - mlp0)[clear], -0.01 * np.sign(w_mlp[clear]), atol=1e-6)
    gclear = np.abs(w_grid) > 0.05 * np.abs(w_grid).max()
    assert np.allclose((grid1 - grid0)[gclear], -0.01 * np.sign(w_grid[gclear]), atol=1e-6)
    assert np.all(grid1[w_grid == 0] == grid0[w_grid == 0])


def test_empty_and_tiny_batches(cache):
    assert cache.train(np.zeros((0, 16), np.float32), np.zeros((0, 3), np.float32)) == 0.0
    rec = world_records(1, 16)
    out = cache.query(rec)
    mlp, grid = cache.get_params()
    want = R.query(rec, stored_grid(grid), mlp, BMIN, inv_extent())
    assert np.allclose(out, want, rtol=3e-2, atol=1e-3 * max(1.0, float(np.abs(want).max())))
    cache.train(rec, np.ones((1, 3), np.float32))


def test_online_fit_reduces_error(ctx):
    """a smooth radiance field over the box is learned from fresh random batches"""
    c = vk.RadianceCache(ctx, vk.default_rc_params(log2_table=16), seed=2)
    c.set_bounds(BMIN, BMAX)

    def field(rec):
        p = (rec[:, 0:3] - BMIN) / (BMAX - BMIN)
        base = 0.5 + 0.4 * np.sin(3.0 * p[:, 0]) * np.cos(2.0 * p[:, 1]) + 0.3 * p[:, 2]
        return np.stack([base, 0.8 * base, 0.5 + 0.2 * rec[:, 5]], axis=1).astype(np.float32)

    test = world_records(4096, 99)
    err0 = float(np.sqrt(np.mean((c.query(test) - field(test)) ** 2)))
    for step in range(300):
        rec = world_records(4096, 1000 + step)
        c.train(rec, field(rec))
    err = float(np.sqrt(np.mean((c.query(test) - field(test)) ** 2)))
    assert err < 0.1 * err0 and err < 0.05, (err0, err)
    c.close()
