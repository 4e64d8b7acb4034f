"""Two-level scenes (instances of mesh BVHs with per-instance transforms) against the oracle's brute force
over every instance (orc_scene_closest): bit-exact records, including instance and custom indices."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from parity import bits
from vierkant_amd import workloads as W

pytestmark = pytest.mark.gpu


def rotation(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis /= np.linalg.norm(axis)
    x, y, z = axis
    c, s = np.cos(angle), np.sin(angle)
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])


def random_instances(n, n_meshes, seed, spread=6.0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        m = np.zeros((3, 4))
        m[:, :3] = rotation(rng.normal(size=3), rng.uniform(0, 2 * np.pi)) @ np.diag(rng.uniform(0.3, 1.5, 3))
        m[:, 3] = rng.uniform(-spread, spread, 3)
        out.append(vk.instance(m, int(rng.integers(n_meshes)), custom_index=1000 + k))
    return np.array(out, dtype=vk.INSTANCE_DTYPE)


@pytest.fixture(scope="module")
def meshes():
    p1, i1 = vk.icosphere(1.0, 2)
    p2, i2 = W.jittered_icosphere(3, seed=4)
    _, i2 = vk.icosphere(1.0, 3)
    p3, i3 = vk.plane(2.0, 2.0, 6, 6)
    return [(p1, i1), (p2.astype(np.float32), i2), (p3, i3)]


def scene_rays(n, seed, spread=8.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    target = rng.uniform(-spread / 2, spread / 2, (n, 3)).astype(np.float32)
    return W.make_rays(o, W.glm_normalize(target - o))


def assert_scene_hits_equal(got, want, tag):
    for f in ("prim", "instance", "custom_index"):
        bad = np.nonzero(got[f] != want[f])[0]
        assert len(bad) == 0, f"{tag}: {f} differs at {bad[:5]}"
    for f in ("t", "u", "v"):
        bad = np.nonzero(bits(got[f]) != bits(want[f]))[0]
        assert len(bad) == 0, f"{tag}: {f} bits differ at {bad[:5]}"


def test_scene_matches_brute_force_over_instances(ctx, meshes):
    bvhs = [ctx.bvh_build(p, i) for p, i in meshes]
    tris = [W.flat_triangles(p, i) for p, i in meshes]
    inst = random_instances(60, len(meshes), seed=1)
    sc = vk.Scene(ctx, bvhs, inst)
    n_inst, n_nodes, stack = sc.info()
    assert n_inst == 60 and n_nodes > 0 and 0 < stack <= 96
    rays = scene_rays(3000, 2)
    got = sc.trace_closest(rays)
    want = orc.scene_closest(tris, inst, rays)
    assert (want["prim"] != 0xFFFFFFFF).sum() > 500
    assert_scene_hits_equal(got, want, "scene closest")
    occ = sc.trace_any(rays)
    assert np.array_equal(occ, (want["prim"] != 0xFFFFFFFF).astype(np.uint8))
    # move every instance: the rebuilt top level stays exact
    inst2 = random_instances(60, len(meshes), seed=7)
    sc.set_instances(inst2)
    assert_scene_hits_equal(sc.trace_closest(rays), orc.scene_closest(tris, inst2, rays), "after set_instances")
    sc.close()
    for b in bvhs:
        b.close()


@pytest.mark.parametrize("n_tri", [1, 2, 8])
def test_single_leaf_meshes(ctx, meshes, n_tri):
    """meshes of at most 8 triangles are one leaf with no node: their bounds (which place the instance in the
    top level) are the padded box of the leaf's triangles, and the scene finds them like any other mesh"""
    p, i = meshes[2]
    small_i = np.ascontiguousarray(i[:3 * n_tri])
    b = ctx.bvh_build(p, small_i, max_leaf_size=max(2, n_tri))
    used = p[small_i].astype(np.float32)
    lo, hi = used.min(axis=0), used.max(axis=0)
    pad = np.float32(2.0 ** -16)
    info = b.info()
    assert info.n_leaves == 1
    assert np.array_equal(np.asarray(info.bounds[:3], np.float32), lo - np.abs(lo) * pad)
    assert np.array_equal(np.asarray(info.bounds[3:], np.float32), hi + np.abs(hi) * pad)
    big = ctx.bvh_build(*meshes[0])
    inst = random_instances(40, 2, seed=11, spread=3.0)
    sc = vk.Scene(ctx, [big, b], inst)
    rays = scene_rays(4000, 12, spread=5.0)
    got = sc.trace_closest(rays)
    want = orc.scene_closest([W.flat_triangles(*meshes[0]), W.flat_triangles(p, small_i)], inst, rays)
    small_hits = (want["prim"] != 0xFFFFFFFF) & (inst["blas"][np.minimum(want["instance"], 39)] == 1)
    assert small_hits.sum() > 5
    assert_scene_hits_equal(got, want, f"{n_tri}-triangle mesh")
    sc.close()
    b.close()
    big.close()


def test_rays_through_many_instances(ctx, meshes):
    """a line of 48 overlapping spheres (instances of two meshes, every fourth one a duplicate of its
    neighbour's placement) and rays along and across it: each ray enters many instance boxes, so the walk
    takes several top-level passes of four kept instances; the records equal the instance brute force, the
    any-hit flags agree, and duplicates tie to the lower instance"""
    bvhs = [ctx.bvh_build(*meshes[0]), ctx.bvh_build(*meshes[1])]
    tris = [W.flat_triangles(*meshes[0]), W.flat_triangles(*meshes[1])]
    inst = []
    for k in range(48):
        j = k - (k % 4 == 3)  # every fourth placement repeats the previous one, mesh included
        m = np.hstack([np.eye(3) * 0.5, np.array([[0.6 * j], [0.0], [0.0]])])
        inst.append(vk.instance(m, j % 2, custom_index=k))
    inst = np.array(inst, dtype=vk.INSTANCE_DTYPE)
    sc = vk.Scene(ctx, bvhs, inst)
    rng = np.random.default_rng(8)
    n = 3000
    o = np.column_stack([np.full(n, -3.0), rng.uniform(-0.4, 0.4, n), rng.uniform(-0.4, 0.4, n)]).astype(np.float32)
    tgt = np.column_stack([np.full(n, 40.0), rng.uniform(-0.4, 0.4, n), rng.uniform(-0.4, 0.4, n)]).astype(np.float32)
    side = np.column_stack([rng.uniform(-1, 29, n), np.full(n, 3.0), rng.uniform(-0.3, 0.3, n)]).astype(np.float32)
    stgt = side + np.float32([0.3, -6.0, 0.0])
    rays = np.concatenate([W.make_rays(o, W.glm_normalize(tgt - o)), W.make_rays(side, W.glm_normalize(stgt - side))])
    # rays that start inside the line: the nearest boxes are behind or around the origin
    inner = W.make_rays(np.column_stack([rng.uniform(0, 28, n), np.zeros(n), np.zeros(n)]).astype(np.float32),
                        W.glm_normalize(rng.normal(size=(n, 3)).astype(np.float32)), tmin=1e-3)
    rays = np.concatenate([rays, inner])
    got = sc.trace_closest(rays)
    want = orc.scene_closest(tris, inst, rays)
    assert (want["prim"] != 0xFFFFFFFF).sum() > 2 * n
    assert_scene_hits_equal(got, want, "many instances")
    assert np.array_equal(sc.trace_any(rays), (want["prim"] != 0xFFFFFFFF).astype(np.uint8))
    hit = got["prim"] != 0xFFFFFFFF
    assert not np.any(hit & (got["instance"] % 4 == 3))  # a repeated placement never beats its original
    sc.close()
    for b in bvhs:
        b.close()


def test_scene_overlapping_copies_tie_to_lower_instance(ctx, meshes):
    p, i = meshes[0]
    b = ctx.bvh_build(p, i)
    ident = np.hstack([np.eye(3), np.zeros((3, 1))])
    inst = np.array([vk.instance(ident, 0, 7), vk.instance(ident, 0, 8)], dtype=vk.INSTANCE_DTYPE)
    sc = vk.Scene(ctx, [b], inst)
    rays = scene_rays(1000, 3, spread=3.0)
    got = sc.trace_closest(rays)
    want = orc.scene_closest([W.flat_triangles(p, i)], inst, rays)
    assert_scene_hits_equal(got, want, "copies")
    hit = got["prim"] != 0xFFFFFFFF
    assert hit.any() and (got["instance"][hit] == 0).all() and (got["custom_index"][hit] == 7).all()
    sc.close()
    b.close()


def test_scene_edge_cases(ctx, meshes):
    rays = scene_rays(200, 4)
    empty_mesh = ctx.bvh_build(np.zeros((3, 3), np.float32), np.zeros(0, np.uint32))
    sc = vk.Scene(ctx, [empty_mesh], np.array([vk.instance(np.hstack([np.eye(3), np.zeros((3, 1))]), 0)],
                                              dtype=vk.INSTANCE_DTYPE))
    got = sc.trace_closest(rays)
    assert (got["prim"] == 0xFFFFFFFF).all() and np.isinf(got["t"]).all()
    assert not sc.trace_any(rays).any()
    sc.close()
    p, i = meshes[2]
    b = ctx.bvh_build(p, i)
    sc = vk.Scene(ctx, [b], np.zeros(0, vk.INSTANCE_DTYPE))
    assert (sc.trace_closest(rays)["prim"] == 0xFFFFFFFF).all()
    bad = np.array([vk.instance(np.hstack([np.eye(3), np.zeros((3, 1))]), 3)], dtype=vk.INSTANCE_DTYPE)
    with pytest.raises(vk.VkrtError):
        sc.set_instances(bad)
    binary = ctx.bvh_build(p, i, width=2)
    with pytest.raises(vk.VkrtError):
        vk.Scene(ctx, [binary], np.zeros(0, vk.INSTANCE_DTYPE))
    sc.close()
    b.close()
    binary.close()
    empty_mesh.close()
