"""GPU interior media (KHR_materials_volume absorption): shadow-ray optical depth
(shaders/slang/ray/visibility_test.slang:110-129) and path segments inside a volume
(raypipeline.slang:380-415, :493) against the CPU integrator and closed forms."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from vierkant_amd import workloads as W
from test_gpu_pt import settings
from test_media_oracle import absorbing_sphere, down_rays, slab_scene

pytestmark = pytest.mark.gpu

# the integrator's expf / logf (glibc) and the GPU's (ocml) may differ in the last ulp
EXP_RTOL = 4e-6


@pytest.mark.parametrize("instanced", [False, True])
def test_slab_transmittance_matches_oracle_and_closed_form(ctx, instanced):
    color, distance = (0.2, 0.6, 0.9), 0.7
    sigma = -np.log(np.asarray(color, np.float64)) / distance
    sc = slab_scene([(0.0, -1.0), (-2.0, -2.5)], color, distance, transmission=0.5, tint=(0.9, 0.8, 0.7))
    per_face = 0.5 * np.array([0.9, 0.8, 0.7])
    pt = vk.PathTracer(ctx, settings(8, 8, 1), seed=0)
    if instanced:
        mf, inst = W.object_meshes(sc)
        pt.set_instanced_scene(sc, mf, inst)
    else:
        pt.set_scene(sc)
    for z0, tmax, want in ((1.0, 10.0, per_face ** 4 * np.exp(-sigma * 1.5)),
                           (-0.25, 10.0, per_face ** 3 * np.exp(-sigma * 1.25)),
                           (1.0, 3.2, per_face ** 3)):
        rays = down_rays(64, z0, tmax)
        got = pt.trace_transmittance(rays)
        ref = orc.pt_transmittance(sc, rays)[:, :3]
        assert np.allclose(got, ref, rtol=EXP_RTOL, atol=0), (z0, tmax)
        assert np.allclose(got, want[None, :], rtol=2e-5, atol=0), (z0, tmax)
    pt.close()


@pytest.mark.parametrize("color", [(0.5, 0.5, 0.5), (0.3, 0.6, 0.9)])
def test_absorbing_sphere_matches_oracle(ctx, color):
    """the index-matched absorbing sphere through its centre, 16,384 samples: the same estimate as the CPU
    integrator (same RNG streams) and the closed form horizon sky * exp(-sigma_c 2r) within the Monte-Carlo
    bound"""
    distance = 1.5
    sc = absorbing_sphere(color, distance)
    cam = vk.make_camera(eye=(0.0, 0.0, 5.0), forward=(0.0, 0.0, -1.0), fov_y=1e-4)
    spp = 16384
    pt = vk.PathTracer(ctx, settings(1, 1, spp, depth=4), seed=0)
    pt.set_scene(sc)
    pt.render(cam)
    got = pt.pixels()["radiance"][0, 0].astype(np.float64)
    pt.close()
    ref = orc.pt_render(sc, cam, 1, 1, spp, 4, W.minstd_rand0(0, 1)[0])[0, 0, :3].astype(np.float64)
    assert np.allclose(got, ref, rtol=1e-3), (got, ref)
    sigma = -np.log(np.asarray(color, np.float64)) / distance
    tr = np.exp(-sigma * 2.0 * 0.9995)
    p = tr.mean()
    sd = 0.6 * (tr / p) * np.sqrt(p * (1 - p) / spp)
    assert np.all(np.abs(got - 0.6 * tr) <= 4.5 * sd + 2e-3 * 0.6 * tr), (got, 0.6 * tr)


def _absorbing_room(scene):
    """the room with every transmissive sphere material given an interior medium (a different tint each)"""
    sc = dict(scene)
    mats = scene["materials"].copy()
    glass = np.flatnonzero(mats["transmission"] > 0)
    assert len(glass) >= 3
    rng = np.random.default_rng(9)
    mats["attenuation_color"][glass] = rng.uniform(0.2, 0.9, (len(glass), 3))
    mats["attenuation_distance"][glass] = rng.uniform(0.1, 0.5, len(glass))
    sc["materials"] = mats
    return sc


@pytest.mark.parametrize("lit", [False, True])
def test_absorbing_room_matches_cpu_integrator(ctx, lit):
    """the room with absorbing glass spheres (paths crossing the media, shadow rays through them), 160x90,
    4 spp, depth 8: the same share-within-1e-4 and RMSE bar as the plain room"""
    sc = _absorbing_room(W.room_scene())
    w, h, spp = 160, 90, 4
    pt = vk.PathTracer(ctx, settings(w, h, spp), seed=0)
    pt.set_scene(sc)
    lights = W.room_lights() if lit else None
    if lit:
        pt.set_lights(lights)
    cam = W.room_camera()
    pt.render(cam)
    got = pt.pixels()
    pt.close()
    want = orc.pt_render(sc, cam, w, h, spp, 8, W.minstd_rand0(0, 1)[0], lights=lights)
    g, c = got["radiance"], want[..., :3]
    assert np.isfinite(g).all()
    assert np.array_equal(got["object_id"], want[..., 6].view(np.uint32))
    diff = np.abs(g - c).max(axis=2)
    close = np.mean(diff <= 1e-4 * np.maximum(1.0, c.max(axis=2)))
    rmse = float(np.sqrt(np.mean((g - c) ** 2)))
    print(f"absorbing room: {close:.4f} of pixels within 1e-4, relative rmse {rmse / float(c.mean()):.5f}")
    assert close > 0.97 and rmse < 0.02 * float(c.mean()) + 1e-6
    # the media change the image against the clear glass
    plain = orc.pt_render(W.room_scene(), cam, w, h, spp, 8, W.minstd_rand0(0, 1)[0], lights=lights)
    assert np.abs(plain[..., :3] - c).max() > 0.05


def test_media_inputs_are_checked(ctx):
    sc = absorbing_sphere((0.5, 0.5, 0.5), 1.0)
    pt = vk.PathTracer(ctx, settings(8, 8, 1), seed=0)
    for color, distance in (((0.0, 0.5, 0.5), 1.0), ((1.5, 0.5, 0.5), 1.0), ((0.5, 0.5, 0.5), -1.0),
                            ((0.5, 0.5, 0.5), float("inf"))):
        bad = dict(sc)
        mats = sc["materials"].copy()
        mats["attenuation_color"][0], mats["attenuation_distance"][0] = color, distance
        bad["materials"] = mats
        with pytest.raises(vk.VkrtError):
            pt.set_scene(bad)
    pt.close()
