"""GPU next-event estimation and MIS (ray/direct_lighting.slang, raypipeline.slang:201-222, :615-619,
:728-761): parity with the CPU integrator on a lit room, the closed-form irradiance cases and the estimator
modes of the oracle's own pins, light validation, and the radiance cache with lights."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from vierkant_amd import workloads as W
from test_nee_oracle import _cases, body_lights, down_camera, plane_scene

pytestmark = pytest.mark.gpu


def settings(w, h, spp, depth=8, mis_mode=0, suppress=0):
    s = vk.default_settings()
    s.resolution[0], s.resolution[1] = w, h
    s.num_samples, s.max_trace_depth = spp, depth
    s.mis_mode, s.suppress_refractive_caustics = mis_mode, suppress
    return s


def render(ctx, scene, lights, cam, w, h, spp, depth=8, seed=0, **kw):
    pt = vk.PathTracer(ctx, settings(w, h, spp, depth, **kw), seed=seed)
    pt.set_scene(scene)
    pt.set_lights(lights)
    pt.render(cam)
    px = pt.pixels()
    pt.close()
    return px


@pytest.mark.parametrize("mode,suppress", [(vk.MIS_FULL, 0), (vk.MIS_NEE_ONLY, 0), (vk.MIS_BSDF_ONLY, 0),
                                           (vk.MIS_FULL, 1)])
def test_lit_room_matches_cpu_integrator(ctx, mode, suppress):
    """room scene with six analytic lights, 160x90, 4 spp, depth 8: per-pixel radiance equals the CPU
    integrator up to transcendental ulps (which can redirect a few paths) and the order the next-event term
    joins the sample sum"""
    sc, cam, lights = W.room_scene(), W.room_camera(), W.room_lights()
    w, h, spp = 160, 90, 4
    got = render(ctx, sc, lights, cam, w, h, spp, mis_mode=mode, suppress=suppress)
    want = orc.pt_render(sc, cam, w, h, spp, 8, W.minstd_rand0(0, 1)[0], lights=lights, mis_mode=mode,
                         suppress_refractive_caustics=suppress)
    g, c = got["radiance"], want[..., :3]
    assert np.isfinite(g).all()
    assert np.array_equal(got["alpha"], want[..., 3])
    diff = np.abs(g - c).max(axis=2)
    close = np.mean(diff <= 1e-4 * np.maximum(1.0, c.max(axis=2)))
    rmse = float(np.sqrt(np.mean((g - c) ** 2)))
    assert close > 0.97, f"only {close:.4f} of pixels within 1e-4"
    assert rmse < 0.02 * float(c.mean()) + 1e-6, f"rmse {rmse} vs mean {c.mean()}"


def test_lights_change_the_image_and_clear(ctx):
    """lights add energy; clearing the list gives back the unlit image bit for bit (same seed)"""
    sc, cam = W.room_scene(), W.room_camera()
    unlit = render(ctx, sc, [], cam, 64, 36, 2)["radiance"]
    lit = render(ctx, sc, W.room_lights(), cam, 64, 36, 2)["radiance"]
    pt = vk.PathTracer(ctx, settings(64, 36, 2), seed=0)
    pt.set_scene(sc)
    pt.set_lights(W.room_lights())
    pt.set_lights([])
    pt.render(cam)
    again = pt.pixels()["radiance"]
    pt.close()
    assert lit.mean() > 1.1 * unlit.mean()
    assert np.array_equal(again, unlit)


@pytest.mark.parametrize("name", list(_cases()))
def test_gpu_nee_closed_form_irradiance(ctx, name):
    """each light type over the rough plane: f * E within 1 % at 8192 samples per pixel, in MIS and NEE-only
    modes"""
    light, want = _cases()[name]
    for mode in (vk.MIS_FULL, vk.MIS_NEE_ONLY):
        px = render(ctx, plane_scene(), [light], down_camera(), 8, 8, 8192, depth=2, seed=1, mis_mode=mode)
        got = float(px["radiance"].mean())
        assert abs(got / want - 1) < 0.01, (name, mode, got, want)


def test_gpu_plane_cases_match_oracle(ctx):
    """the closed-form setups pixel by pixel against the CPU integrator at 64 spp"""
    lights = np.array([c[0] for c in _cases().values()], vk.LIGHT_DTYPE)
    px = render(ctx, plane_scene(), lights, down_camera(), 16, 16, 64, depth=3, seed=2)
    want = orc.pt_render(plane_scene(), down_camera(), 16, 16, 64, 3, W.minstd_rand0(2, 1)[0], lights=lights)
    g, c = px["radiance"], want[..., :3]
    close = np.mean(np.abs(g - c).max(axis=2) <= 1e-4 * np.maximum(1.0, c.max(axis=2)))
    assert close > 0.97, close


def test_gpu_estimator_modes_agree(ctx):
    """single-lobe glossy metal under every light body kind: NEE only, BSDF only and MIS agree within noise
    at 16384 samples per pixel"""
    sc = plane_scene(roughness=0.4, metalness=1.0)
    means = {}
    for mode in (vk.MIS_FULL, vk.MIS_NEE_ONLY, vk.MIS_BSDF_ONLY):
        means[mode] = float(render(ctx, sc, body_lights(), down_camera(), 8, 8, 16384, depth=2, seed=3,
                                   mis_mode=mode)["radiance"].mean())
    full = means[vk.MIS_FULL]
    assert abs(means[vk.MIS_NEE_ONLY] / full - 1) < 0.01 and abs(means[vk.MIS_BSDF_ONLY] / full - 1) < 0.02, means


def test_set_lights_validation(ctx):
    pt = vk.PathTracer(ctx, settings(8, 8, 1), seed=0)
    bad = vk.light(vk.LIGHT_SPHERE, intensity=1.0, size=(0.0, 0.0))
    with pytest.raises(vk.VkrtError):
        pt.set_lights([bad])
    unknown = vk.light(vk.LIGHT_OMNI, intensity=1.0)
    unknown["type"] = 9
    with pytest.raises(vk.VkrtError):
        pt.set_lights([unknown])
    s = settings(8, 8, 1)
    s.mis_mode = 3
    with pytest.raises(vk.VkrtError):
        pt.settings = s
    pt.close()


def test_radiance_cache_with_lights_converges_near_plain(ctx):
    """the cache's training targets include next-event estimates and the light bodies met on each segment:
    after training on the lit room its image is within 6 % of plain path tracing's mean"""
    sc, cam, lights = W.room_scene(), W.room_camera(), W.room_lights()
    w, h = 96, 54
    ref = render(ctx, sc, lights, cam, w, h, 256, seed=5)["radiance"]
    rc = vk.RadianceCache(ctx, seed=1)
    pt = vk.PathTracer(ctx, settings(w, h, 4), seed=6)
    pt.set_scene(sc)
    pt.set_lights(lights)
    pt.set_radiance_cache(rc, cache_depth=2, train_stride=8)
    for _ in range(64):
        pt.render(cam)
    got = pt.pixels()["radiance"]
    pt.close()
    rc.close()
    assert np.isfinite(got).all()
    assert abs(got.mean() / ref.mean() - 1) < 0.06, (got.mean(), ref.mean())
