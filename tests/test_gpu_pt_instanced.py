"""GPU path tracer over two-level scenes (PathTracer.set_instanced_scene): bottom-level BVHs per mesh under a
top-level BVH of instances, object rays through each instance's world-to-object rows, shading points,
normals, tangents and texture footprints through its transform, tmax inside a volume from its scaled box
(src/RayBuilder.cpp:386-781, shaders/slang/raypipeline.slang:299-370)."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from vierkant_amd import workloads as W
from test_gpu_pt import _alpha_room, settings
from test_pt_oracle_instanced import _glassy, _rays_in_room, glass_quad_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room():
    return W.room_scene()


@pytest.fixture(scope="module")
def room_inst():
    sc, mf, inst = W.room_instanced()
    return _glassy(sc, mf), mf, inst


@pytest.mark.parametrize("alpha", [False, True])
@pytest.mark.parametrize("lit", [False, True])
def test_identity_instances_render_the_flat_image(ctx, room, lit, alpha):
    """the room (276,504 triangles) split into one mesh per object, each placed once by an identity
    transform: the same image bit for bit as the flat scene and the same segment count (object rays equal the
    world rays, (t, instance, triangle) order equals (t, triangle) order; lit: the six analytic lights and
    their shadow rays; alpha: blended and masked materials, the any-hit filter in every walk)"""
    sc = _alpha_room(room) if alpha else room
    mf, inst = W.object_meshes(sc)
    out = []
    for two_level in (False, True):
        pt = vk.PathTracer(ctx, settings(320, 180, 2), seed=6)
        if two_level:
            pt.set_instanced_scene(sc, mf, inst)
        else:
            pt.set_scene(sc)
        if lit:
            pt.set_lights(W.room_lights())
        pt.render(W.room_camera())
        out.append((pt.pixels(), pt.stats().segments))
        pt.close()
    assert out[0][0].tobytes() == out[1][0].tobytes()
    assert out[0][1] == out[1][1] > 320 * 180 * 2


@pytest.mark.parametrize("lit", [False, True])
def test_instanced_room_matches_cpu_integrator(ctx, room_inst, lit):
    """the instanced room (200 placements of 8 shared unit spheres under random rotations and scales, half
    of them glass, a rotated non-uniformly scaled icosphere), 160x90, 4 spp, depth 8: per-pixel radiance
    equals the CPU reference integrator's two-level scene up to transcendental ulps, as for flat scenes"""
    sc, mf, inst = room_inst
    w, h, spp = 160, 90, 4
    pt = vk.PathTracer(ctx, settings(w, h, spp), seed=0)
    pt.set_instanced_scene(sc, mf, inst)
    lights = W.room_lights() if lit else None
    if lit:
        pt.set_lights(lights)
    cam = W.room_camera()
    pt.render(cam)
    got = pt.pixels()
    pt.close()
    want = orc.pt_render(sc, cam, w, h, spp, 8, W.minstd_rand0(0, 1)[0], lights=lights, instanced=(mf, inst))
    g, c = got["radiance"], want[..., :3]
    assert np.isfinite(g).all()
    assert np.array_equal(got["alpha"], want[..., 3])
    assert np.array_equal(got["object_id"], want[..., 6].view(np.uint32))
    diff = np.abs(g - c).max(axis=2)
    close = np.mean(diff <= 1e-4 * np.maximum(1.0, c.max(axis=2)))
    rmse = float(np.sqrt(np.mean((g - c) ** 2)))
    print(f"instanced GPU vs integrator: {close:.4f} of pixels within 1e-4, relative rmse {rmse / float(c.mean()):.5f}")
    assert close > 0.97, f"only {close:.4f} of pixels within 1e-4"
    assert rmse < 0.02 * float(c.mean()) + 1e-6, f"rmse {rmse} vs mean {c.mean()}"


def test_scaled_glass_square_matches_cpu_integrator(ctx):
    """a glass square scaled by 2 above the floor: refracted rays run to tmax = 2 |scale * half extents|
    and reach the floor, as in the CPU integrator"""
    sc, mf, cam = glass_quad_scene()
    q = (0.0, float(np.sin(np.pi / 4)), 0.0, float(np.cos(np.pi / 4)))
    inst = vk.pt_instances([0, 1], [(0, 0, 0), (0, 2.0, 0)], [(0, 0, 0, 1), q], [(1, 1, 1), (2, 2, 2)])
    pt = vk.PathTracer(ctx, settings(32, 32, 4, depth=4), seed=0)
    pt.set_instanced_scene(sc, mf, inst)
    pt.render(cam)
    got = pt.pixels()["radiance"]
    pt.close()
    want = orc.pt_render(sc, cam, 32, 32, 4, 4, W.minstd_rand0(0, 1)[0], instanced=(mf, inst))[..., :3]
    assert np.mean(np.abs(got - want).max(axis=2) <= 1e-4 * np.maximum(1.0, want.max(axis=2))) > 0.97
    assert (got[12:20, 12:20, 0] > 0.3).mean() > 0.9


@pytest.mark.parametrize("cull", [False, True])
def test_instanced_transmittance_matches_oracle(ctx, room_inst, cull):
    """transmittance_test through the instanced room: every record identical to the oracle's two-level
    query (candidates in (t, instance, triangle) order, back faces judged in object space)"""
    sc, mf, inst = room_inst
    rays = _rays_in_room(2000, 8)
    pt = vk.PathTracer(ctx, settings(8, 8, 1), seed=0)
    pt.set_instanced_scene(sc, mf, inst)
    got = pt.trace_transmittance(rays, cull_back=cull)
    pt.close()
    want = orc.pt_transmittance(sc, rays, cull=cull, instanced=(mf, inst))[:, :3]
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert (got == 0).any(1).sum() > 100 and ((got > 0) & (got < 1)).all(1).sum() > 10


def test_cache_inactive_below_its_depth_is_plain_pt_on_instances(ctx, room_inst):
    """the cache mode's training paths walk the two-level scene too: with the query depth at
    max_trace_depth the image is plain path tracing bit for bit"""
    sc, mf, inst = room_inst
    plain = vk.PathTracer(ctx, settings(96, 54, 2), seed=4)
    cached = vk.PathTracer(ctx, settings(96, 54, 2), seed=4)
    cache = vk.RadianceCache(ctx, vk.default_rc_params(log2_table=16), seed=1)
    for pt in (plain, cached):
        pt.set_instanced_scene(sc, mf, inst)
        pt.set_lights(W.room_lights())
    cached.set_radiance_cache(cache, cache_depth=8, train_stride=4)
    for pt in (plain, cached):
        pt.render(W.room_camera())
    assert plain.pixels().tobytes() == cached.pixels().tobytes()
    assert cache.step() == 2
    cache.close()


def test_scene_kinds_replace_each_other(ctx, room):
    """a flat scene set after an instanced one (and the reverse) renders as on a fresh tracer; rejected
    instanced scenes (ranges that do not cover the triangles, a mesh index out of range, a zero scale, a
    non-finite rotation) leave the previous scene in place; no instances renders the sky (each tracer's first
    frame is compared: later frames draw other samples)"""
    mf, inst = W.object_meshes(room)
    cam = W.room_camera()

    def first_image(setup):
        pt = vk.PathTracer(ctx, settings(64, 36, 1), seed=2)
        setup(pt)
        pt.render(cam)
        px = pt.pixels()
        pt.close()
        return px

    want = first_image(lambda pt: pt.set_scene(room)).tobytes()

    def flat_after_instanced(pt):
        pt.set_instanced_scene(room, mf, inst)
        pt.set_scene(room)

    assert first_image(flat_after_instanced).tobytes() == want
    bad = [(mf[:-1], inst), (mf, vk.pt_instances([len(mf)])),
           (mf, vk.pt_instances([0], scales=[(1.0, 0.0, 1.0)])),
           (mf, vk.pt_instances([0], rotations=[(np.nan, 0.0, 0.0, 1.0)]))]

    def instanced_after_flat_and_rejections(pt):
        pt.set_scene(room)
        pt.set_instanced_scene(room, mf, inst)
        for m, i in bad:
            with pytest.raises(vk.VkrtError):
                pt.set_instanced_scene(room, m, i)

    assert first_image(instanced_after_flat_and_rejections).tobytes() == want
    sky = first_image(lambda pt: pt.set_instanced_scene(room, mf, vk.pt_instances([])))
    assert (sky["alpha"] == 0).all() and np.isfinite(sky["radiance"]).all()
