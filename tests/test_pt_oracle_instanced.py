"""Pins the CPU reference integrator's two-level scenes (oracle/vkpt_oracle.c, the instance level of
src/RayBuilder.cpp:386-781 with the shading transforms of shaders/slang/raypipeline.slang:299-370):
identity instances reproduce the flat scene bit for bit, a placed instance renders like its triangles moved
into world space, the instance matrix is the textbook rotation matrix, and the transmittance query agrees
with the brute-force transmittance oracle."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from vierkant_amd import workloads as W


def _alpha(scene):
    sc = dict(scene)
    mats = scene["materials"].copy()
    mats["alpha_mode"][0::3] = vk.ALPHA_BLEND
    mats["color"][0::3, 3] = 0.5
    mats["alpha_mode"][1::3] = vk.ALPHA_MASK
    mats["color"][1::3, 3] = 0.3
    mats["alpha_cutoff"][1::3] = 0.5
    sc["materials"] = mats
    return sc


@pytest.mark.parametrize("alpha", [False, True])
@pytest.mark.parametrize("lit", [False, True])
def test_identity_instances_render_the_flat_scene(lit, alpha):
    """the room split into one mesh per object, each placed once by an identity transform: the object rays,
    shading points and normals equal the flat scene's, and (t, instance, triangle) order equals (t, triangle)
    order, so the image is the flat one bit for bit"""
    sc = W.room_scene(n_spheres=20)
    sc = _alpha(sc) if alpha else sc
    mf, inst = W.object_meshes(sc)
    assert len(mf) - 1 == len(inst) > 20
    lights = W.room_lights() if lit else None
    cam = W.room_camera()
    flat = orc.pt_render(sc, cam, 48, 27, 2, 6, 7, lights=lights)
    two = orc.pt_render(sc, cam, 48, 27, 2, 6, 7, lights=lights, instanced=(mf, inst))
    assert flat.tobytes() == two.tobytes()
    assert (flat[..., 3] > 0).mean() > 0.25


def test_duplicate_instances_render_like_one():
    """every mesh placed twice at the same spot: equal t, the lower instance wins and shades the same
    point, so the image is unchanged"""
    sc = W.room_scene(n_spheres=10)
    mf, inst = W.object_meshes(sc)
    dup = np.concatenate([inst, inst])
    cam = W.room_camera()
    a = orc.pt_render(sc, cam, 40, 24, 2, 6, 3, instanced=(mf, inst))
    b = orc.pt_render(sc, cam, 40, 24, 2, 6, 3, instanced=(mf, dup))
    assert a.tobytes() == b.tobytes()


def _rotation_matrix(q):
    x, y, z, w = (float(c) for c in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_instance_matrix_is_rotation_times_scale():
    """pt_instance_matrix (the library's object-to-world matrix of a vkrt_pt_instance) against the textbook
    quaternion rotation matrix times diag(scale), translation in the last column"""
    rng = np.random.default_rng(4)
    q = W._unit_quaternions(rng.uniform(size=(50, 3)))
    s = rng.uniform(0.2, 3.0, (50, 3)).astype(np.float32)
    t = rng.uniform(-5, 5, (50, 3)).astype(np.float32)
    inst = vk.pt_instances(np.zeros(50), t, q, s)
    for rec in inst:
        m = orc.pt_instance_matrix(rec).astype(np.float64)
        want = _rotation_matrix(rec["rotation"]) * rec["scale"].astype(np.float64)[None, :]
        assert np.allclose(m[:, :3], want, atol=2e-6 * float(rec["scale"].max())), rec
        assert np.array_equal(m[:, 3], rec["translation"].astype(np.float64))
    # a half turn about y scales and flips exactly
    m = orc.pt_instance_matrix(vk.pt_instances([0], [(1, 2, 3)], [(0, 1, 0, 0)], [(2, 2, 2)])[0])
    assert np.array_equal(m, np.array([[-2, 0, 0, 1], [0, 2, 0, 2], [0, 0, -2, 3]], np.float32))


def _world_scene(scene, mesh_first, inst):
    """the two-level scene with every instance's triangles moved into world space (apply_transform in double,
    normals turned by the rotation only, as the shading does), one object per instance and mesh object"""
    pos, nrm, idx, tob, omat = [], [], [], [], []
    nv = 0
    for k, rec in enumerate(inst):
        a, b = int(mesh_first[rec["mesh"]]), int(mesh_first[rec["mesh"] + 1])
        tri = scene["indices"].reshape(-1, 3)[a:b]
        used, inv = np.unique(tri.reshape(-1), return_inverse=True)
        m = orc.pt_instance_matrix(rec).astype(np.float64)
        p = scene["positions"][used].astype(np.float64)
        pos.append((p @ m[:, :3].T + m[:, 3]).astype(np.float32))
        r = _rotation_matrix(rec["rotation"])
        nrm.append((scene["normals"][used].astype(np.float64) @ r.T).astype(np.float32))
        idx.append(inv.astype(np.uint32) + nv)
        nv += len(used)
        objs = scene["tri_object"][a:b]
        base = len(omat)
        uo, oinv = np.unique(objs, return_inverse=True)
        tob.append((oinv + base).astype(np.uint32))
        omat.extend(scene["object_material"][uo])
    return {"positions": np.concatenate(pos), "normals": np.concatenate(nrm), "indices": np.concatenate(idx),
            "tri_object": np.concatenate(tob), "object_material": np.asarray(omat, np.uint32),
            "materials": scene["materials"], "environment": scene["environment"]}


def _glassy(scene, mesh_first):
    """the instanced room with every other shared sphere mesh (1, 3, 5, ...) made of glass (transmission 1,
    IOR 1.5), so shadow rays multiply through instances and refracted rays take the scaled tmax inside"""
    sc = dict(scene)
    mats = scene["materials"].copy()
    for m in range(1, len(mesh_first) - 2, 2):
        k = scene["object_material"][scene["tri_object"][mesh_first[m]]]
        mats["transmission"][k], mats["ior"][k], mats["metalness"][k] = 1.0, 1.5, 0.0
    sc["materials"] = mats
    return sc


def test_placed_instances_render_like_world_space_triangles():
    """the instanced room (shared unit spheres under random rotations and scales, a rotated, non-uniformly
    scaled icosphere): rendered through object rays it matches the same triangles moved into world space up
    to rounding of the two arithmetics (a transposed matrix, a conjugated quaternion or a missing scale moves
    whole objects and fails this)"""
    sc, mf, inst = W.room_instanced(n_spheres=40)
    sc = _glassy(sc, mf)
    flat = _world_scene(sc, mf, inst)
    cam = W.room_camera()
    a = orc.pt_render(sc, cam, 64, 36, 4, 6, 11, instanced=(mf, inst))
    b = orc.pt_render(flat, cam, 64, 36, 4, 6, 11)
    assert np.array_equal(a[..., 3], b[..., 3])
    cov = a[..., 3] > 0
    assert np.allclose(a[..., 4][cov], b[..., 4][cov], rtol=1e-4, atol=1e-4)
    diff = np.abs(a[..., :3] - b[..., :3]).max(axis=2)
    close = np.mean(diff <= 1e-3 * np.maximum(1.0, b[..., :3].max(axis=2)))
    assert close > 0.9, close
    assert abs(float(a[..., :3].mean()) - float(b[..., :3].mean())) < 0.02 * float(b[..., :3].mean())


def glass_quad_scene():
    """mesh 0: an 8 x 8 white floor at y = 0; mesh 1: a unit glass square (transmission 1, IOR 1) at y = 0,
    both facing +y; the camera looks straight down from y = 6"""
    pos = np.array([(-4, 0, -4), (4, 0, -4), (4, 0, 4), (-4, 0, 4),
                    (-0.5, 0, -0.5), (0.5, 0, -0.5), (0.5, 0, 0.5), (-0.5, 0, 0.5)], np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (8, 1))
    idx = np.array([0, 2, 1, 0, 3, 2, 4, 6, 5, 4, 7, 6], np.uint32)
    mats = np.array([vk.material((0.8, 0.8, 0.8), roughness=1.0),
                     vk.material((1, 1, 1), roughness=0.0, transmission=1.0, ior=1.0)], vk.MATERIAL_DTYPE)
    sc = {"positions": pos, "normals": nrm, "indices": idx, "tri_object": np.array([0, 0, 1, 1], np.uint32),
          "object_material": np.array([0, 1], np.uint32), "materials": mats, "environment": 1.0}
    cam = vk.make_camera(eye=(0.0, 6.0, 0.0), forward=(0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0), fov_y=0.6)
    return sc, np.array([0, 2, 4], np.uint32), cam


def test_scaled_instance_limits_refracted_rays_by_its_scaled_box():
    """tmax_inside_volume = 2 |scale * half extents| (raypipeline.slang:370): the glass square scaled by 2
    (and turned a quarter about y) 2 units above the floor lets refracted rays reach the floor (tmax 2.83),
    as the same square built 2 x 2 in world space does; with its unscaled box (tmax 1.41) they would stop
    short and see the sky"""
    sc, mf, cam = glass_quad_scene()
    q = (0.0, float(np.sin(np.pi / 4)), 0.0, float(np.cos(np.pi / 4)))
    inst = vk.pt_instances([0, 1], [(0, 0, 0), (0, 2.0, 0)], [(0, 0, 0, 1), q], [(1, 1, 1), (2, 2, 2)])
    a = orc.pt_render(sc, cam, 32, 32, 4, 4, 1, instanced=(mf, inst))
    b = orc.pt_render(_world_scene(sc, mf, inst), cam, 32, 32, 4, 4, 1)
    diff = np.abs(a[..., :3] - b[..., :3]).max(axis=2)
    assert np.mean(diff <= 1e-3) > 0.95
    # below the square the floor (about 0.5-1.6) is seen, not the sky below the horizon (0.1)
    assert (a[12:20, 12:20, 0] > 0.3).mean() > 0.9


def _rays_in_room(n, seed):
    rng = np.random.default_rng(seed)
    o = np.column_stack([rng.uniform(-2.8, 2.8, n), rng.uniform(0.2, 5.8, n), rng.uniform(-2.8, 2.8, n)])
    p = np.column_stack([rng.uniform(-2.8, 2.8, n), rng.uniform(0.2, 5.8, n), rng.uniform(-2.8, 2.8, n)])
    d = p - o
    dist = np.linalg.norm(d, axis=1)
    return W.make_rays(o.astype(np.float32), W.glm_normalize((d / dist[:, None]).astype(np.float32)), tmin=1e-4,
                       tmax=dist.astype(np.float32))


@pytest.mark.parametrize("cull", [False, True])
def test_scene_transmittance_matches_brute_force(cull):
    """the integrator's transmittance query on the flat room equals the brute-force transmittance oracle bit
    for bit, and the identity-instanced room equals the flat one"""
    sc = W.room_scene(n_spheres=40)
    rays = _rays_in_room(600, 2)
    want = orc.transmittance(sc, rays, cull=cull)
    got = orc.pt_transmittance(sc, rays, cull=cull)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    two = orc.pt_transmittance(sc, rays, cull=cull, instanced=W.object_meshes(sc))
    assert np.array_equal(two.view(np.uint32), want.view(np.uint32))
    assert (want[:, 0] == 0).sum() > 50 and ((want[:, :3] > 0) & (want[:, :3] < 1)).all(1).sum() > 5


def test_placed_transmittance_matches_world_space_triangles():
    sc, mf, inst = W.room_instanced(n_spheres=40)
    sc = _glassy(sc, mf)
    flat = _world_scene(sc, mf, inst)
    rays = _rays_in_room(600, 3)
    a = orc.pt_transmittance(sc, rays, instanced=(mf, inst))
    b = orc.pt_transmittance(flat, rays)
    same = np.all(np.abs(a - b) <= 1e-5, axis=1)
    assert same.mean() > 0.99, same.mean()
    assert ((b[:, :3] > 0) & (b[:, :3] < 1)).all(1).sum() > 5
