"""The PLOC builder (agglomerative clustering over the Morton order): traversal through its tree equals the
brute-force loop over the reference test bit for bit, like the LBVH's, on the same cases, and its surface
area cost on the 1.8M-triangle scene is below the LBVH's."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from parity import assert_hits_equal
from vierkant_amd import workloads as W

pytestmark = pytest.mark.gpu
PLOC = vk.BUILDER_PLOC


@pytest.fixture(scope="module")
def icosphere_case():
    pos, idx, rays = W.icosphere_rays()
    tris = W.flat_triangles(pos, idx)
    return pos, idx, tris, rays, orc.brute_closest(tris, rays)


@pytest.mark.parametrize("leaf", [1, 2, 4, 8])
@pytest.mark.parametrize("width", [0, 2])
def test_ploc_closest_icosphere_bit_exact(ctx, icosphere_case, leaf, width):
    pos, idx, tris, rays, want = icosphere_case
    bvh = ctx.bvh_build(pos, idx, builder=PLOC, max_leaf_size=leaf, width=width)
    info = bvh.info()
    assert info.n_tri == len(idx) // 3 and 1 < info.depth <= 64
    assert_hits_equal(ctx.trace_closest(bvh, rays), want, f"ploc leaf={leaf} width={width}")
    bvh.close()


def test_ploc_any_and_windowed(ctx, icosphere_case):
    pos, idx, tris, rays, _ = icosphere_case
    bvh = ctx.bvh_build(pos, idx, builder=PLOC)
    r = rays.copy()
    rng = np.random.default_rng(30)
    r["tmax"] = rng.uniform(0.0, 6.0, len(r)).astype(np.float32)
    r["tmin"] = rng.uniform(-0.5, 0.5, len(r)).astype(np.float32)
    assert np.array_equal(ctx.trace_any(bvh, r), orc.brute_any(tris, r))
    assert_hits_equal(ctx.trace_closest(bvh, r), orc.brute_closest(tris, r), "ploc windowed")


@pytest.mark.parametrize("n_tri", [0, 1, 2, 3, 7, 9, 33, 100, 1000])
def test_ploc_small_meshes_and_duplicates(ctx, n_tri):
    rng = np.random.default_rng(31 + n_tri)
    pos = rng.uniform(-1, 1, (max(3 * n_tri, 1), 3)).astype(np.float32)
    idx = np.arange(3 * n_tri, dtype=np.uint32)
    if n_tri >= 9:
        idx[3 * (n_tri // 2):3 * (n_tri // 2) + 9] = idx[:9]   # three repeated triangles: equal boxes
    tris = W.flat_triangles(pos, idx) if n_tri else np.zeros((0, 9), np.float32)
    o = rng.uniform(-3, 3, (3000, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(3000, 3)).astype(np.float32))
    rays = W.make_rays(o, d)
    for leaf in (1, 4):
        bvh = ctx.bvh_build(pos, idx, builder=PLOC, max_leaf_size=leaf)
        assert_hits_equal(ctx.trace_closest(bvh, rays), orc.brute_closest(tris, rays), f"n_tri={n_tri}")
        assert np.array_equal(ctx.trace_any(bvh, rays), orc.brute_any(tris, rays))
        bvh.close()


def test_ploc_degenerate_nonfinite_and_refit(ctx, icosphere_case):
    rng = np.random.default_rng(32)
    pos = rng.uniform(-1, 1, (600, 3)).astype(np.float32)
    pos[0:3] = pos[0]
    pos[9] = [np.inf, 0, 0]
    pos[12] = [np.nan, 0, 0]
    idx = np.arange(600, dtype=np.uint32)
    tris = W.flat_triangles(pos, idx)
    o = rng.uniform(-3, 3, (4000, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(4000, 3)).astype(np.float32))
    rays = W.make_rays(o, d)
    bvh = ctx.bvh_build(pos, idx, builder=PLOC, max_leaf_size=2)
    assert_hits_equal(ctx.trace_closest(bvh, rays), orc.brute_closest(tris, rays), "ploc degenerate")
    pos2, idx2, _, rays2, _ = icosphere_case
    b2 = ctx.bvh_build(pos2, idx2, builder=PLOC)
    moved = (pos2 * np.float32(0.7) - np.float32(0.2)).astype(np.float32)
    b2.refit(moved)
    assert_hits_equal(ctx.trace_closest(b2, rays2[:8192]),
                      orc.brute_closest(W.flat_triangles(moved, idx2), rays2[:8192]), "ploc refit")


def test_ploc_heightfield_matches_and_lowers_sah(ctx):
    """1,804,288 triangles: primary and bounce rays through the PLOC tree equal GPU brute force and the LBVH
    tree's records; its SAH cost is lower"""
    pos, idx = W.heightfield_spheres_scene()
    tris = W.flat_triangles(pos, idx)
    ploc = ctx.bvh_build(pos, idx, builder=PLOC)
    lbvh = ctx.bvh_build(pos, idx)
    pi, li = ploc.info(), lbvh.info()
    assert pi.sah_cost < 0.95 * li.sah_cost, (pi.sah_cost, li.sah_cost)
    assert pi.depth <= 64 and pi.n_wide_nodes > 0
    primary_all = W.pinhole_rays(2048, 2048)
    sel = np.random.default_rng(33).choice(len(primary_all), 65536, replace=False)
    primary = primary_all[np.sort(sel)]
    hits = ctx.trace_closest(ploc, primary)
    assert_hits_equal(hits, ctx.brute_closest(tris, primary), "ploc heightfield primary")
    assert_hits_equal(hits, ctx.trace_closest(lbvh, primary), "ploc vs lbvh primary")
    bounce, shadow = W.secondary_rays(primary, hits, tris)
    assert_hits_equal(ctx.trace_closest(ploc, bounce), ctx.brute_closest(tris, bounce), "ploc heightfield bounce")
    assert np.array_equal(ctx.trace_any(ploc, shadow), ctx.brute_any(tris, shadow))
