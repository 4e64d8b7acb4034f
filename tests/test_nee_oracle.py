"""Pins the next-event estimation of the CPU integrator (oracle/vkpt_oracle.c): analytic lights over a rough
plane seen straight down, against closed-form irradiance (form factors and inverse-square law, worked out
here from the geometry), the three estimator modes against each other, and light bodies seen directly."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk

RHO = 0.5
# rough dielectric at normal incidence and view: Disney diffuse rho / pi (Fd = 1), GGX specular with
# alpha = 1: F = 0.08, D = 1 / pi, visibility 1 / 4, so f = (rho + 0.02) / pi
F_NORMAL = (RHO + 0.02) / np.pi
H_LIGHT = 10.0


def plane_scene(rho=RHO, roughness=1.0, metalness=0.0, half=50.0):
    """a quad at y = 0 facing +y, no sky"""
    pos = np.array([[-half, 0, -half], [half, 0, -half], [half, 0, half], [-half, 0, half]], np.float32)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    idx = np.array([0, 2, 1, 0, 3, 2], np.uint32)
    m = vk.material(color=(rho, rho, rho), roughness=roughness, metalness=metalness)
    return {"positions": pos, "normals": nrm, "indices": idx, "tri_object": np.zeros(2, np.uint32),
            "object_material": np.zeros(1, np.uint32), "materials": np.array([m], vk.MATERIAL_DTYPE),
            "environment": 0.0}


def down_camera():
    """5 units above the plane looking straight down, a 4 degree field: every pixel sees the plane at
    nearly normal incidence, within 0.2 units of the origin"""
    return vk.make_camera(eye=(0.0, 5.0, 0.0), forward=(0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0),
                          fov_y=float(np.radians(4.0)))


def render_mean(lights, spp, mode=vk.MIS_FULL, scene=None, seed=1, size=8):
    img = orc.pt_render(scene or plane_scene(), down_camera(), size, size, spp, 2, seed, lights=lights,
                        mis_mode=mode)
    return img[..., :3].mean(), img


def corner_form_factor(a, b, c):
    """point to a parallel a x b rectangle at distance c, the point below one corner (Howell's catalogue)"""
    A, B = a / c, b / c
    sa, sb = np.sqrt(1 + A * A), np.sqrt(1 + B * B)
    return (A / sa * np.arctan(B / sa) + B / sb * np.arctan(A / sb)) / (2 * np.pi)


def tube_irradiance(radius, half_len, height, le, n=1200):
    """cosine-weighted irradiance at the origin (normal +y) from a horizontal tube along x at `height`,
    by midpoint quadrature over its mantle, front-facing part only"""
    phi = (np.arange(n) + 0.5) * 2 * np.pi / n
    z = (np.arange(n) + 0.5) * 2 * half_len / n - half_len
    P, Z = np.meshgrid(phi, z, indexing="ij")
    nq = np.stack([np.zeros_like(P), np.cos(P), np.sin(P)], -1)
    q = np.stack([Z, height + radius * np.cos(P), radius * np.sin(P)], -1)
    d2 = (q * q).sum(-1)
    d = np.sqrt(d2)
    cos_p = q[..., 1] / d
    cos_l = -(nq * q).sum(-1) / d
    dA = radius * (2 * np.pi / n) * (2 * half_len / n)
    return le * np.sum(np.where(cos_l > 0, cos_p * cos_l / d2, 0.0) * dA)


def _cases():
    le = 100.0
    r = 1.0
    sin2 = r * r / (H_LIGHT * H_LIGHT)
    ff = 4 * corner_form_factor(1.0, 0.5, H_LIGHT)
    sun = 0.1
    return {
        "sphere": (vk.light(vk.LIGHT_SPHERE, position=(0, H_LIGHT, 0), intensity=le, size=(r, 0)),
                   F_NORMAL * le * np.pi * sin2),
        "disk": (vk.light(vk.LIGHT_DISK, position=(0, H_LIGHT, 0), direction=(0, -1, 0), intensity=le, size=(r, 0)),
                 F_NORMAL * le * np.pi * r * r / (H_LIGHT ** 2 + r * r)),
        "rect": (vk.light(vk.LIGHT_RECT, position=(0, H_LIGHT, 0), direction=(0, -1, 0), tangent=(1, 0, 0),
                          intensity=le, size=(1.0, 0.5)), F_NORMAL * le * np.pi * ff),
        "tube": (vk.light(vk.LIGHT_TUBE, position=(0, H_LIGHT, 0), direction=(1, 0, 0), intensity=le,
                          size=(0.2, 1.0)), F_NORMAL * tube_irradiance(0.2, 1.0, H_LIGHT, le)),
        "omni": (vk.light(vk.LIGHT_OMNI, position=(0, H_LIGHT, 0), intensity=50.0), F_NORMAL * 50.0 / H_LIGHT ** 2),
        "spot": (vk.light(vk.LIGHT_SPOT, position=(0, H_LIGHT, 0), direction=(0, -1, 0), intensity=50.0,
                          inner_cone_angle=0.1, outer_cone_angle=0.5), F_NORMAL * 50.0 / H_LIGHT ** 2),
        "directional": (vk.light(vk.LIGHT_DIRECTIONAL, direction=(0, -1, 0), intensity=2.0), F_NORMAL * 2.0),
        "sun_disc": (vk.light(vk.LIGHT_DIRECTIONAL, direction=(0, -1, 0), intensity=le, angular_size=sun),
                     F_NORMAL * le * np.pi * np.sin(sun) ** 2),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_nee_matches_closed_form_irradiance(name):
    """f * E for each light type, by next-event estimation alone and with MIS (the BSDF rays from the plane
    meet the light body in the ghost-light pass): within 1.5 % at 1024 samples per pixel"""
    light, want = _cases()[name]
    for mode in (vk.MIS_NEE_ONLY, vk.MIS_FULL):
        got, img = render_mean([light], 1024, mode)
        assert abs(got / want - 1) < 0.015, (name, mode, got, want)
        assert np.isfinite(img).all()


def body_lights():
    """one light of every kind BSDF rays can reach, large and tilted"""
    lights = [vk.light(vk.LIGHT_SPHERE, position=(3.0, 6.0, 1.0), intensity=20.0, size=(2.5, 0)),
              vk.light(vk.LIGHT_RECT, position=(-4.0, 5.0, 0.0), direction=(0.6, -0.8, 0.0), tangent=(0.0, 0.0, 1.0),
                       intensity=15.0, size=(2.0, 2.0)),
              vk.light(vk.LIGHT_DISK, position=(0.0, 4.0, -4.0), direction=(0.0, -0.6, 0.8), intensity=15.0,
                       size=(2.0, 0)),
              vk.light(vk.LIGHT_TUBE, position=(0.0, 4.0, 3.0), direction=(1.0, 0.0, 0.0), intensity=15.0,
                       size=(0.6, 3.0)),
              vk.light(vk.LIGHT_DIRECTIONAL, direction=(0.3, -0.9, 0.3), intensity=10.0, angular_size=0.4)]
    for k in (1, 4):
        lights[k]["direction"] /= np.linalg.norm(lights[k]["direction"])
    return lights


def test_estimator_modes_agree_on_a_tilted_glossy_setup():
    """the reference's convergence check (ray_common.slang:288-294): NEE only, BSDF only and MIS estimate the
    same integral for every light BSDF rays can reach; large lights keep the BSDF-only estimate tight. The
    material is a single-lobe glossy metal: the hit-side MIS weight takes the pdf of the lobe sample_disney
    drew (bsdf_disney.slang:162-296) while the NEE side takes eval_disney's mixture pdf, so for several lobes
    the reference's modes differ by design (see test_mis_weights_follow_the_sampled_lobe)"""
    sc = plane_scene(roughness=0.4, metalness=1.0)
    lights = body_lights()
    full, _ = render_mean(lights, 2048, vk.MIS_FULL, sc)
    nee, _ = render_mean(lights, 2048, vk.MIS_NEE_ONLY, sc)
    bsdf, _ = render_mean(lights, 2048, vk.MIS_BSDF_ONLY, sc)
    assert abs(nee / full - 1) < 0.02 and abs(bsdf / full - 1) < 0.04, (full, nee, bsdf)


def test_light_body_seen_directly_is_its_radiance():
    """a camera ray that meets a sphere light adds color * intensity with weight 1 (no NEE competitor at the
    first segment), bit for bit; NEE-only mode leaves light bodies black to BSDF rays"""
    le = np.float32(37.5)
    light = vk.light(vk.LIGHT_SPHERE, position=(0, 10, 0), color=(1.0, 0.5, 0.25), intensity=le, size=(2.0, 0))
    cam = vk.make_camera(eye=(0.0, 5.0, 0.0), forward=(0.0, 1.0, 0.0), up=(0.0, 0.0, -1.0),
                         fov_y=float(np.radians(4.0)))
    img = orc.pt_render(plane_scene(), cam, 6, 6, 4, 3, 2, lights=[light])
    want = np.array([1.0, 0.5, 0.25], np.float32) * le
    assert np.array_equal(img[..., :3], np.broadcast_to(want, img[..., :3].shape))
    img = orc.pt_render(plane_scene(), cam, 6, 6, 4, 3, 2, lights=[light], mis_mode=vk.MIS_NEE_ONLY)
    assert (img[..., :3] == 0).all()


def test_lights_are_gathered_like_the_reference():
    """zero-intensity and Area lights are dropped before the pick (PBRPathTracer.cpp:630-637): adding them
    changes nothing, bit for bit"""
    sc = plane_scene()
    base = [vk.light(vk.LIGHT_SPHERE, position=(1, 8, 0), intensity=30.0, size=(1.0, 0)),
            vk.light(vk.LIGHT_OMNI, position=(-1, 6, 1), intensity=20.0)]
    extra = [base[0], vk.light(vk.LIGHT_OMNI, position=(0, 3, 0), intensity=0.0), base[1],
             vk.light(vk.LIGHT_AREA, position=(0, 3, 0), intensity=5.0)]
    a = orc.pt_render(sc, down_camera(), 8, 8, 16, 3, 4, lights=base)
    b = orc.pt_render(sc, down_camera(), 8, 8, 16, 3, 4, lights=extra)
    assert a.tobytes() == b.tobytes()


def test_mis_weights_follow_the_sampled_lobe():
    """with a diffuse + specular mixture the hit-side weight uses the drawn lobe's pdf and the NEE side the
    mixture pdf, as the reference does; the two single-strategy modes still agree with each other, and MIS
    lands below them (the weights of a direction sum to less than one)"""
    sc = plane_scene(roughness=0.4, metalness=0.5)
    lights = body_lights()
    full, _ = render_mean(lights, 2048, vk.MIS_FULL, sc)
    nee, _ = render_mean(lights, 2048, vk.MIS_NEE_ONLY, sc)
    bsdf, _ = render_mean(lights, 2048, vk.MIS_BSDF_ONLY, sc)
    assert abs(bsdf / nee - 1) < 0.04 and full < 0.93 * nee, (full, nee, bsdf)
