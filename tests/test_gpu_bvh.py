"""GPU parity of BVH traversal: closest and any hit through the BVH equal the brute-force loop over the
reference test (bit-exact records, lowest index on ties), for every builder setting, edge-case meshes,
refit, and the 1.8M-triangle heightfield scene."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from parity import assert_hits_equal
from vierkant_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def icosphere_case():
    pos, idx, rays = W.icosphere_rays()
    tris = W.flat_triangles(pos, idx)
    return pos, idx, tris, rays, orc.brute_closest(tris, rays)


@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_bvh_closest_icosphere_bit_exact(ctx, icosphere_case, leaf):
    pos, idx, tris, rays, want = icosphere_case
    bvh = ctx.bvh_build(pos, idx, max_leaf_size=leaf)
    info = bvh.info()
    assert info.n_tri == len(idx) // 3 and 1 < info.depth <= 64
    got = ctx.trace_closest(bvh, rays)
    assert_hits_equal(got, want, f"bvh leaf={leaf}")


def test_bvh_any_icosphere(ctx, icosphere_case):
    pos, idx, tris, rays, _ = icosphere_case
    bvh = ctx.bvh_build(pos, idx)
    r = rays.copy()
    rng = np.random.default_rng(20)
    r["tmax"] = rng.uniform(0.0, 6.0, len(r)).astype(np.float32)
    r["tmin"] = rng.uniform(-0.5, 0.5, len(r)).astype(np.float32)
    want = orc.brute_any(tris, r)
    got = ctx.trace_any(bvh, r)
    assert np.array_equal(got, want)
    # the same rays, closest hit with the interval [tmin, tmax]
    assert_hits_equal(ctx.trace_closest(bvh, r), orc.brute_closest(tris, r), "bvh closest windowed")


def test_bvh_duplicate_triangles_tie_break(ctx):
    rng = np.random.default_rng(21)
    pos = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    idx = rng.integers(0, 300, (150, 3)).astype(np.uint32)
    idx = np.concatenate([idx, idx[::-1], idx]).reshape(-1)      # each triangle three times
    tris = W.flat_triangles(pos, idx)
    o = rng.uniform(-3, 3, (5000, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(5000, 3)).astype(np.float32))
    rays = W.make_rays(o, d)
    want = orc.brute_closest(tris, rays)
    assert np.sum(want["prim"] != 0xFFFFFFFF) > 200
    for leaf in (1, 3, 8):
        bvh = ctx.bvh_build(pos, idx, max_leaf_size=leaf)
        assert_hits_equal(ctx.trace_closest(bvh, rays), want, f"dups leaf={leaf}")


@pytest.mark.parametrize("n_tri", [0, 1, 2, 3, 7, 9, 100])
def test_bvh_small_meshes(ctx, n_tri):
    rng = np.random.default_rng(22 + n_tri)
    pos = rng.uniform(-1, 1, (max(3 * n_tri, 1), 3)).astype(np.float32)
    idx = np.arange(3 * n_tri, dtype=np.uint32)
    tris = W.flat_triangles(pos, idx) if n_tri else np.zeros((0, 9), np.float32)
    o = rng.uniform(-3, 3, (2000, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(2000, 3)).astype(np.float32))
    rays = W.make_rays(o, d)
    bvh = ctx.bvh_build(pos, idx)
    assert_hits_equal(ctx.trace_closest(bvh, rays), orc.brute_closest(tris, rays), f"n_tri={n_tri}")
    assert np.array_equal(ctx.trace_any(bvh, rays), orc.brute_any(tris, rays))


def test_bvh_degenerate_and_nonfinite(ctx):
    rng = np.random.default_rng(23)
    pos = rng.uniform(-1, 1, (600, 3)).astype(np.float32)
    pos[0:3] = pos[0]                       # a point triangle
    pos[3:6, 1] = 0.5                       # a flat triangle in y
    pos[9] = [np.inf, 0, 0]
    pos[12] = [np.nan, 0, 0]
    idx = np.arange(600, dtype=np.uint32)
    tris = W.flat_triangles(pos, idx)
    o = rng.uniform(-3, 3, (4000, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(4000, 3)).astype(np.float32))
    d[:10] = [0, 0, 1]                       # axis-aligned rays
    d[10:20] = [0, 1, 0]
    rays = W.make_rays(o, d)
    bvh = ctx.bvh_build(pos, idx, max_leaf_size=2)
    assert_hits_equal(ctx.trace_closest(bvh, rays), orc.brute_closest(tris, rays), "degenerate")


def test_bvh_bad_index_is_an_error(ctx):
    pos = np.zeros((3, 3), np.float32)
    idx = np.array([0, 1, 3], np.uint32)
    with pytest.raises(vk.VkrtError) as e:
        ctx.bvh_build(pos, idx)
    assert e.value.code == vk.E_INVALID_ARG


def test_bvh_stride(ctx, icosphere_case):
    pos, idx, tris, rays, want = icosphere_case
    padded = np.zeros((len(pos), 8), np.float32)
    padded[:, :3] = pos
    bvh = ctx.bvh_build(padded, idx, stride=32)
    assert_hits_equal(ctx.trace_closest(bvh, rays[:4096]), want[:4096], "stride 32")


def test_bvh_refit(ctx, icosphere_case):
    pos, idx, tris, rays, want = icosphere_case
    bvh = ctx.bvh_build(pos, idx)
    moved = (pos * np.float32(1.3) + np.float32(0.1)).astype(np.float32)
    bvh.refit(moved)
    tris2 = W.flat_triangles(moved, idx)
    want2 = orc.brute_closest(tris2, rays[:8192])
    assert_hits_equal(ctx.trace_closest(bvh, rays[:8192]), want2, "refit")


def test_refit_keeps_the_4wide_shape(ctx, icosphere_case):
    """a refit refreshes the boxes of the 4-wide collapse and keeps its shape: the node count, the stack bound
    and the 4-wide flag stay what the build chose (scenes and batches keep them from creation), also under a
    deformation that skews the boxes (half the sphere stretched 50 times along x); traces stay exact"""
    pos, idx, tris, rays, want = icosphere_case
    bvh = ctx.bvh_build(pos, idx)
    before = bvh.info()
    skewed = pos.copy()
    skewed[skewed[:, 0] > 0, 0] *= np.float32(50.0)
    bvh.refit(skewed)
    after = bvh.info()
    assert (after.n_wide_nodes, after.wide_stack) == (before.n_wide_nodes, before.wide_stack)
    assert after.n_wide_nodes > 0
    # the refit's bounds hold the deformed mesh, padded by at most 2^-16 relative; depth and SAH cost stay
    # the build's statistics
    b = np.array(after.bounds[:6], np.float64)
    lo, hi = skewed.min(axis=0).astype(np.float64), skewed.max(axis=0).astype(np.float64)
    assert np.all(b[:3] <= lo) and np.all(b[3:] >= hi)
    assert np.allclose(b, np.concatenate([lo, hi]), rtol=2 ** -15, atol=0)
    assert (after.depth, after.sah_cost) == (before.depth, before.sah_cost)
    want2 = orc.brute_closest(W.flat_triangles(skewed, idx), rays[:8192])
    assert_hits_equal(ctx.trace_closest(bvh, rays[:8192]), want2, "skewed refit")


@pytest.fixture(scope="module")
def heightfield_case():
    pos, idx = W.heightfield_spheres_scene()
    return pos, idx, W.flat_triangles(pos, idx)


def test_bvh_heightfield_primary_and_bounce(ctx, heightfield_case):
    """1,804,288 triangles: 65,536 primary and 65,536 bounce rays through the BVH equal GPU brute force
    (itself bit-exact with the oracle on the 20,480-triangle icosphere), and 1,024 of them equal the oracle directly"""
    pos, idx, tris = heightfield_case
    bvh = ctx.bvh_build(pos, idx)
    primary_all = W.pinhole_rays(2048, 2048)
    sel = np.random.default_rng(24).choice(len(primary_all), 65536, replace=False)
    primary = primary_all[np.sort(sel)]
    hits = ctx.trace_closest(bvh, primary)
    assert np.mean(hits["prim"] != 0xFFFFFFFF) > 0.5
    assert_hits_equal(hits, ctx.brute_closest(tris, primary), "heightfield_case primary")
    bounce, shadow = W.secondary_rays(primary, hits, tris)
    bh = ctx.trace_closest(bvh, bounce)
    assert_hits_equal(bh, ctx.brute_closest(tris, bounce), "heightfield_case bounce")
    assert np.array_equal(ctx.trace_any(bvh, shadow), ctx.brute_any(tris, shadow))
    few = np.arange(0, 65536, 64)
    assert_hits_equal(hits[few], orc.brute_closest(tris, primary[few]), "heightfield_case vs oracle")
    assert_hits_equal(bh[few], orc.brute_closest(tris, bounce[few]), "heightfield_case bounce vs oracle")


def test_trace_configs_give_identical_results(ctx, icosphere_case):
    """every launch shape (block size x LDS stack depth x loop of the 4-wide walk) and both traversal layouts
    return the same records"""
    pos, idx, tris, rays, want = icosphere_case
    saved = ctx.get_trace_config()
    try:
        for width in (0, 2):
            bvh = ctx.bvh_build(pos, idx, width=width)
            for block, lds, trav, order in [(b, l, t, o) for b in (64, 128, 256) for l in (0, 8, 16, 32)
                                            for t in (vk.TRAVERSAL_STEP, vk.TRAVERSAL_POSTPONED)
                                            for o in (0, 1, 16)]:
                    ctx.set_trace_config(block, lds, trav, order)
                    tag = f"width={width} block={block} lds={lds} traversal={trav} xcd_chunk={order}"
                    assert_hits_equal(ctx.trace_closest(bvh, rays), want, tag)
                    occ = ctx.trace_any(bvh, rays)
                    assert np.array_equal(occ, (want["prim"] != 0xFFFFFFFF).astype(np.uint8)), tag
            bvh.close()
        # ragged ray counts under XCD chunks: partial last blocks and whole idle rounds of workgroups
        bvh = ctx.bvh_build(pos, idx)
        for n in (1, 63, 1000, 16385):
            for chunk in (0, 1, 32, 256):
                ctx.set_trace_config(64, 16, vk.TRAVERSAL_STEP, chunk)
                assert_hits_equal(ctx.trace_closest(bvh, rays[:n]), want[:n], f"n={n} xcd_chunk={chunk}")
        bvh.close()
        with pytest.raises(vk.VkrtError):
            ctx.set_trace_config(96, 16)
        with pytest.raises(vk.VkrtError):
            ctx.set_trace_config(256, 16, 2)
        with pytest.raises(vk.VkrtError):
            ctx.set_trace_config(256, 16, 0, 3)
    finally:
        ctx.set_trace_config(*saved)


@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_wide_collapse_equals_binary_tree(ctx, heightfield_case, leaf):
    """the 4-wide collapse (default layout) and the binary tree give identical closest and any hits on
    primary and bounce rays of the heightfield + spheres scene; the collapse has about a third as many nodes"""
    pos, idx, tris = heightfield_case
    primary = W.pinhole_rays(256, 256)
    wide = ctx.bvh_build(pos, idx, max_leaf_size=leaf)
    binary = ctx.bvh_build(pos, idx, max_leaf_size=leaf, width=2)
    wi, bi = wide.info(), binary.info()
    assert wi.n_wide_nodes > 0 and bi.n_wide_nodes == 0
    assert wi.n_wide_nodes < 0.5 * wi.n_nodes and 0 < wi.wide_stack <= 64
    hw = ctx.trace_closest(wide, primary)
    hb = ctx.trace_closest(binary, primary)
    assert_hits_equal(hw, hb, "wide vs binary primary")
    bounce, shadow = W.secondary_rays(primary, hb, tris)
    assert_hits_equal(ctx.trace_closest(wide, bounce), ctx.trace_closest(binary, bounce), "wide vs binary bounce")
    assert np.array_equal(ctx.trace_any(wide, shadow), ctx.trace_any(binary, shadow))
    sel = np.arange(0, len(bounce), 97)
    assert_hits_equal(ctx.trace_closest(wide, bounce[sel]), ctx.brute_closest(tris, bounce[sel]), "wide vs brute")
    # both loops of the 4-wide walk and the workgroup orders give the same records
    saved = ctx.get_trace_config()
    try:
        for trav in (vk.TRAVERSAL_STEP, vk.TRAVERSAL_POSTPONED):
            for order in (0, 4, 256):
                ctx.set_trace_config(saved[0], saved[1], trav, order)
                tag = f"traversal={trav} xcd_chunk={order}"
                assert_hits_equal(ctx.trace_closest(wide, primary), hb, tag + " primary")
                assert_hits_equal(ctx.trace_closest(wide, bounce), ctx.trace_closest(binary, bounce), tag + " bounce")
                assert np.array_equal(ctx.trace_any(wide, shadow), ctx.trace_any(binary, shadow)), tag
    finally:
        ctx.set_trace_config(*saved)
    wide.close()
    binary.close()
