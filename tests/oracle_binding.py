"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference-compiled helpers in
oracle/_ref/libvkref.so. Test infrastructure: only tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py use it, and only as the comparison reference."""
import ctypes
import os
import pathlib

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle" / "liboracle.so"
REF = ROOT / "oracle" / "_ref" / "libvkref.so"

_vp, _u32, _i32, _f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float
_fp, _dp = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)

_SIG = {
    "orc_intersect_tri": (_i32, [_vp, _vp, _vp, _fp, _fp, _fp]),
    "orc_intersect_tri_f64": (_i32, [_vp, _vp, _vp, _dp, _dp, _dp]),
    "orc_intersect_obb": (_i32, [_vp, _vp, _vp, _fp]),
    "orc_intersect_aabb": (_i32, [_vp, _vp, _vp, _fp]),
    "orc_intersect_sphere": (_i32, [_vp, _vp, _vp, _fp]),
    "orc_intersect_plane": (_i32, [_vp, _vp, _vp, _fp]),
    "orc_compute_aabb": (None, [_vp, _u32, _vp]),
    "orc_bounding_sphere": (None, [_vp, _u32, _vp]),
    "orc_centroid": (None, [_vp, _u32, _vp]),
    "orc_scene_closest": (None, [_vp, _vp, _vp, _vp, _vp, _u32, _vp, _u32, _vp]),
    "orc_denoise": (None, [_vp, _u32, _u32, ctypes.c_int, _vp]),
    "orc_half_value": (ctypes.c_float, [ctypes.c_float]),
    "orc_tonemap": (None, [_vp, ctypes.c_size_t, ctypes.c_float, ctypes.c_float, _vp]),
    "orc_transmittance": (None, [_vp, _vp, _vp, _vp, _u32, _vp, _u32, ctypes.c_int, _vp]),
    "orc_brute_closest": (None, [_vp, _u32, _vp, _u32, _vp, _i32]),
    "orc_brute_any": (None, [_vp, _u32, _vp, _u32, _vp, _i32]),
    "orc_brute_closest_cull_safe": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _i32]),
    "orc_cull_safe": (_i32, [_vp, _vp, _f32]),
    "orc_closest_f64": (_u32, [_vp, _u32, _vp, _dp, _dp, _dp]),
    "orc_icosphere_size": (None, [_u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "orc_icosphere": (None, [_f32, _u32, _vp, _vp]),
    "orc_plane": (None, [_f32, _f32, _u32, _u32, _vp, _vp]),
    "orc_xxhash32": (_u32, [_u32, _u32]),
    "orc_lcg": (_u32, [ctypes.POINTER(_u32)]),
    "orc_rnd": (_f32, [ctypes.POINTER(_u32)]),
    "orc_hammersley": (None, [_u32, _u32, _vp]),
}

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE.exists():
            raise RuntimeError(f"{ORACLE} missing: run `make -C oracle`")
        h = ctypes.CDLL(str(ORACLE))
        for name, (res, args) in _SIG.items():
            getattr(h, name).restype = res
            getattr(h, name).argtypes = args
        _lib = h
    return _lib


def ref():
    """reference-compiled triangle_intersection.h helpers, or None when oracle/_ref was not built"""
    global _ref
    if _ref is None and REF.exists():
        h = ctypes.CDLL(str(REF))
        h.ref_tri_box_overlap.restype = _u32
        h.ref_tri_box_overlap.argtypes = [_vp, _vp]
        h.ref_tri_tri_overlap.restype = _u32
        h.ref_tri_tri_overlap.argtypes = [_vp, _vp]
        _ref = h
    return _ref


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data


def threads():
    return int(os.environ.get("ORACLE_THREADS", min(16, os.cpu_count() or 1)))


def intersect_tri(tri9, o, d):
    tri9, pt = _f(tri9)
    o, po = _f(o)
    d, pd = _f(d)
    t, u, v = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    r = lib().orc_intersect_tri(pt, po, pd, ctypes.byref(t), ctypes.byref(u), ctypes.byref(v))
    return (r, t.value, u.value, v.value) if r else (0, 0.0, 0.0, 0.0)


def intersect_tri_f64(tri9, o, d):
    tri9, pt = _f(tri9)
    o, po = _f(o)
    d, pd = _f(d)
    t, u, v = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    r = lib().orc_intersect_tri_f64(pt, po, pd, ctypes.byref(t), ctypes.byref(u), ctypes.byref(v))
    return (r, t.value, u.value, v.value) if r else (0, 0.0, 0.0, 0.0)


def _volume(fn, prim, o, d):
    prim, pp = _f(prim)
    o, po = _f(o)
    d, pd = _f(d)
    dist = ctypes.c_float()
    r = fn(pp, po, pd, ctypes.byref(dist))
    return (r, dist.value) if r else (0, 0.0)


def intersect_aabb(box6, o, d):
    return _volume(lib().orc_intersect_aabb, box6, o, d)


def intersect_obb(obb15, o, d):
    return _volume(lib().orc_intersect_obb, obb15, o, d)


def intersect_sphere(s4, o, d):
    return _volume(lib().orc_intersect_sphere, s4, o, d)


def intersect_plane(p4, o, d):
    return _volume(lib().orc_intersect_plane, p4, o, d)


def ray_od(rays):
    o = np.stack([rays["ox"], rays["oy"], rays["oz"]], axis=1).astype(np.float32)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], axis=1).astype(np.float32)
    return o, d


def compute_aabb(points):
    points, pp = _f(points)
    out = np.zeros(6, dtype=np.float32)
    lib().orc_compute_aabb(pp, len(points) // 3 if points.ndim == 1 else len(points), out.ctypes.data)
    return out


def scene_closest(mesh_tris, instances, rays):
    """mesh_tris: list of flat (n, 9) triangle arrays; instances: INSTANCE_DTYPE records; returns
    INSTANCE_HIT_DTYPE records"""
    from vierkant_amd import INSTANCE_HIT_DTYPE
    tris = np.ascontiguousarray(np.concatenate([np.asarray(t, np.float32).reshape(-1, 9) for t in mesh_tris]) if mesh_tris
                                else np.zeros((0, 9), np.float32))
    off = np.zeros(len(mesh_tris) + 1, np.uint32)
    off[1:] = np.cumsum([len(np.asarray(t).reshape(-1, 9)) for t in mesh_tris])
    w2o = np.ascontiguousarray(instances["world_to_object"], np.float32)
    mesh = np.ascontiguousarray(instances["blas"], np.uint32)
    custom = np.ascontiguousarray(instances["custom_index"], np.uint32)
    r = np.ascontiguousarray(rays)
    out = np.zeros(len(r), INSTANCE_HIT_DTYPE)
    lib().orc_scene_closest(tris.ctypes.data, off.ctypes.data, w2o.ctypes.data, mesh.ctypes.data, custom.ctypes.data,
                            len(instances), r.ctypes.data, len(r), out.ctypes.data)
    return out


def transmittance(scene, rays, cull=False):
    """brute-force transmittance_test over a path-tracer scene dict (positions, indices, tri_object,
    object_material, materials)"""
    tris = np.ascontiguousarray(
        scene["positions"][scene["indices"].reshape(-1, 3)].reshape(-1, 9).astype(np.float32))
    mats = scene["materials"][scene["object_material"][scene["tri_object"]]]
    trans = np.ascontiguousarray(mats["transmission"], np.float32)
    color = np.ascontiguousarray(mats["color"][:, :3], np.float32)
    two = np.ascontiguousarray(mats["two_sided"] != 0, np.uint8)
    r = np.ascontiguousarray(rays)
    out = np.zeros((len(r), 4), np.float32)
    lib().orc_transmittance(tris.ctypes.data, trans.ctypes.data, color.ctypes.data, two.ctypes.data, len(tris),
                            r.ctypes.data, len(r), int(cull), out.ctypes.data)
    return out


def denoise(pixels, denoise=True):
    """the a-trous denoise pass over a (H, W) PIXEL_DTYPE buffer"""
    px = np.ascontiguousarray(pixels)
    h, w = px.shape
    out = np.zeros((h, w, 4), np.float32)
    lib().orc_denoise(px.ctypes.data, w, h, int(denoise), out.ctypes.data)
    return out


def denoise_passes(pixels, passes):
    """only the first `passes` a-trous passes (steps 1, 2, 4, 8, 16) over a (H, W) PIXEL_DTYPE buffer"""
    h_ = lib()
    h_.orc_denoise_passes.restype = None
    h_.orc_denoise_passes.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                      ctypes.c_void_p]
    px = np.ascontiguousarray(pixels)
    h, w = px.shape
    out = np.zeros((h, w, 4), np.float32)
    h_.orc_denoise_passes(px.ctypes.data, w, h, int(passes), out.ctypes.data)
    return out


def half_value(x):
    """float32 array rounded to binary16 values (nearest, ties to even) by the C restatement"""
    h = lib()
    return np.array([h.orc_half_value(float(v)) for v in np.asarray(x, np.float32).ravel()], np.float32)


def tonemap(rgba, exposure=2.0, gamma=1.0):
    """the exposure/gamma composite over (..., 4) float32 texels"""
    a = np.ascontiguousarray(rgba, np.float32)
    out = np.empty_like(a)
    lib().orc_tonemap(a.ctypes.data, a.size // 4, exposure, gamma, out.ctypes.data)
    return out


def bounding_sphere(points):
    points, pp = _f(points)
    out = np.zeros(4, dtype=np.float32)
    lib().orc_bounding_sphere(pp, len(points) // 3 if points.ndim == 1 else len(points), out.ctypes.data)
    return out


def centroid(points):
    points, pp = _f(points)
    out = np.zeros(3, dtype=np.float32)
    lib().orc_centroid(pp, len(points) // 3 if points.ndim == 1 else len(points), out.ctypes.data)
    return out


def brute_closest(tris_flat, rays, nthreads=None):
    from vierkant_amd import HIT_DTYPE
    tris_flat, pt = _f(tris_flat)
    rays = np.ascontiguousarray(rays)
    out = np.zeros(len(rays), dtype=HIT_DTYPE)
    lib().orc_brute_closest(pt, tris_flat.size // 9, rays.ctypes.data, len(rays), out.ctypes.data,
                            nthreads or threads())
    return out


def brute_closest_cull_safe(tris_flat, rays, nthreads=None):
    """(hits, safe): the brute-force closest hits and, per ray, 1 when the library's padded box of the winning
    triangle is entered with best t = the winner's t (then no BVH of the library can cull the winner)"""
    from vierkant_amd import HIT_DTYPE
    tris_flat, pt = _f(tris_flat)
    rays = np.ascontiguousarray(rays)
    out = np.zeros(len(rays), dtype=HIT_DTYPE)
    safe = np.zeros(len(rays), dtype=np.uint8)
    lib().orc_brute_closest_cull_safe(pt, tris_flat.size // 9, rays.ctypes.data, len(rays), out.ctypes.data,
                                      safe.ctypes.data, nthreads or threads())
    return out, safe


def cull_safe(tri9, ray, t):
    tri9 = np.ascontiguousarray(tri9, dtype=np.float32)
    ray = np.ascontiguousarray(ray)
    return int(lib().orc_cull_safe(tri9.ctypes.data, ray.ctypes.data, float(t)))


def brute_any(tris_flat, rays, nthreads=None):
    tris_flat, pt = _f(tris_flat)
    rays = np.ascontiguousarray(rays)
    out = np.zeros(len(rays), dtype=np.uint8)
    lib().orc_brute_any(pt, tris_flat.size // 9, rays.ctypes.data, len(rays), out.ctypes.data,
                        nthreads or threads())
    return out


def closest_f64(tris_flat, ray):
    tris_flat, pt = _f(tris_flat)
    ray = np.ascontiguousarray(np.asarray(ray).reshape(1))
    t, u, v = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    prim = lib().orc_closest_f64(pt, tris_flat.size // 9, ray.ctypes.data, ctypes.byref(t), ctypes.byref(u),
                                 ctypes.byref(v))
    return prim, t.value, u.value, v.value


def icosphere(radius, subdiv):
    nv, ni = _u32(), _u32()
    lib().orc_icosphere_size(subdiv, ctypes.byref(nv), ctypes.byref(ni))
    pos = np.zeros((nv.value, 3), dtype=np.float32)
    idx = np.zeros(ni.value, dtype=np.uint32)
    lib().orc_icosphere(radius, subdiv, pos.ctypes.data, idx.ctypes.data)
    return pos, idx


def plane(w, h, sw, sh):
    pos = np.zeros(((sw + 1) * (sh + 1), 3), dtype=np.float32)
    idx = np.zeros(sw * sh * 6, dtype=np.uint32)
    lib().orc_plane(w, h, sw, sh, pos.ctypes.data, idx.ctypes.data)
    return pos, idx


def xxhash32(a, b):
    return lib().orc_xxhash32(a, b)


def rnd_sequence(state, n):
    s = _u32(state)
    return [lib().orc_rnd(ctypes.byref(s)) for _ in range(n)], s.value


def hammersley(i, n):
    out = np.zeros(2, dtype=np.float32)
    lib().orc_hammersley(i, n, out.ctypes.data)
    return out


class _PtScene(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("normals", ctypes.c_void_p), ("n_vert", ctypes.c_uint32),
                ("indices", ctypes.c_void_p), ("n_tri", ctypes.c_uint32), ("tri_object", ctypes.c_void_p),
                ("object_material", ctypes.c_void_p), ("n_objects", ctypes.c_uint32),
                ("materials", ctypes.c_void_p), ("n_materials", ctypes.c_uint32), ("environment", ctypes.c_float),
                ("texcoords", ctypes.c_void_p), ("textures", ctypes.c_void_p), ("n_textures", ctypes.c_uint32),
                ("tangents", ctypes.c_void_p), ("n_meshes", ctypes.c_uint32), ("mesh_first", ctypes.c_void_p),
                ("n_instances", ctypes.c_uint32), ("inst_mesh", ctypes.c_void_p), ("inst_w2o", ctypes.c_void_p),
                ("inst_xf", ctypes.c_void_p)]


def pt_instance_matrix(rec):
    """the 3x4 object-to-world matrix of one vkrt_pt_instance: column j = R(q) (scale_j e_j) as
    v + 2 q.xyz x (q.xyz x v + w v) in double, the last column the translation, rounded once to float"""
    qx, qy, qz, qw = (float(x) for x in rec["rotation"])
    m = np.zeros((3, 4), np.float32)
    for j in range(3):
        v = [0.0, 0.0, 0.0]
        v[j] = float(rec["scale"][j])
        c1 = [qy * v[2] - qz * v[1] + qw * v[0], qz * v[0] - qx * v[2] + qw * v[1], qx * v[1] - qy * v[0] + qw * v[2]]
        c2 = [qy * c1[2] - qz * c1[1], qz * c1[0] - qx * c1[2], qx * c1[1] - qy * c1[0]]
        for a in range(3):
            m[a, j] = np.float32(v[a] + 2.0 * c2[a])
    m[:, 3] = rec["translation"]
    return m


def _pt_instanced(instanced):
    """(mesh_first u32, world-to-object rows f32 [n,12], mesh u32, transform f32 [n,10]) of (mesh_first,
    PT_INSTANCE_DTYPE records); the world-to-object rows are vkrt_instance_set_transform's inverse of
    pt_instance_matrix, the matrices the tracer builds its instances from"""
    import vierkant_amd as vk
    mesh_first, inst = instanced
    mf = np.ascontiguousarray(mesh_first, np.uint32)
    inst = np.asarray(inst, vk.PT_INSTANCE_DTYPE).reshape(-1)
    w2o = np.zeros((max(1, len(inst)), 12), np.float32)
    for k, rec in enumerate(inst):
        w2o[k] = vk.instance(pt_instance_matrix(rec), int(rec["mesh"]))["world_to_object"]
    xf = np.zeros((max(1, len(inst)), 10), np.float32)
    if len(inst):
        xf[:, 0:4], xf[:, 4:7], xf[:, 7:10] = inst["rotation"], inst["translation"], inst["scale"]
    return mf, w2o, np.ascontiguousarray(inst["mesh"], np.uint32), xf


def _pt_scene(scene, instanced=None):
    """an orc_pt_scene over a scene dict (and, for a two-level scene, (mesh_first, instances)), with the
    arrays it points into"""
    arr = {k: np.ascontiguousarray(scene[k]) for k in
           ("positions", "normals", "indices", "tri_object", "object_material", "materials")}
    uv = scene.get("texcoords")
    uv = np.ascontiguousarray(uv, np.float32) if uv is not None and len(uv) else None
    tg = scene.get("tangents")
    tg = np.ascontiguousarray(tg, np.float32) if tg is not None and len(tg) else None
    imgs = [np.ascontiguousarray(t, np.uint8) for t in scene.get("textures", [])]
    smp = list(scene.get("texture_samplers") or [(0, 0, 0, 0)] * len(imgs))
    tex = (_Texture * max(1, len(imgs)))(*[_Texture(t.shape[1], t.shape[0], t.ctypes.data, *s)
                                           for t, s in zip(imgs, smp)])
    ins = _pt_instanced(instanced) if instanced is not None else None
    sc = _PtScene(arr["positions"].ctypes.data, arr["normals"].ctypes.data, len(arr["positions"]),
                  arr["indices"].ctypes.data, len(arr["indices"]) // 3, arr["tri_object"].ctypes.data,
                  arr["object_material"].ctypes.data, len(arr["object_material"]), arr["materials"].ctypes.data,
                  len(arr["materials"]), float(scene.get("environment", 1.0)),
                  uv.ctypes.data if uv is not None else None, ctypes.addressof(tex) if imgs else None, len(imgs),
                  tg.ctypes.data if tg is not None else None)
    if ins is not None:
        mf, w2o, mesh, xf = ins
        sc.n_meshes, sc.mesh_first = len(mf) - 1, mf.ctypes.data
        sc.n_instances, sc.inst_mesh = len(mesh), mesh.ctypes.data
        sc.inst_w2o, sc.inst_xf = w2o.ctypes.data, xf.ctypes.data
    return sc, (arr, uv, tg, imgs, tex, ins)


def pt_transmittance(scene, rays, cull=False, instanced=None):
    """vkrt_pt_trace_transmittance's product per ray (rgb, 0) through the oracle's scene, flat or two-level"""
    h = lib()
    h.orc_pt_transmittance.restype = None
    h.orc_pt_transmittance.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                                       ctypes.c_void_p]
    sc, keep = _pt_scene(scene, instanced)
    r = np.ascontiguousarray(rays)
    out = np.zeros((len(r), 4), np.float32)
    h.orc_pt_transmittance(ctypes.byref(sc), r.ctypes.data, len(r), int(cull), out.ctypes.data)
    del keep
    return out


class _Texture(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("rgba", ctypes.c_void_p),
                ("address_mode_u", ctypes.c_uint32), ("address_mode_v", ctypes.c_uint32),
                ("min_filter", ctypes.c_uint32), ("mag_filter", ctypes.c_uint32)]


class _PtParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("num_samples", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("seed", ctypes.c_uint32), ("disable_material", ctypes.c_uint32),
                ("max_path_beta", ctypes.c_float), ("eye", ctypes.c_float * 3), ("forward", ctypes.c_float * 3),
                ("right", ctypes.c_float * 3), ("up", ctypes.c_float * 3), ("tan_half_x", ctypes.c_float),
                ("tan_half_y", ctypes.c_float), ("aperture", ctypes.c_float), ("focal_distance", ctypes.c_float),
                ("lights", ctypes.c_void_p), ("n_lights", ctypes.c_uint32), ("mis_mode", ctypes.c_uint32),
                ("suppress_refractive_caustics", ctypes.c_uint32), ("fov_y", ctypes.c_float),
                ("test_flags", ctypes.c_uint32)]


PT_NO_ROULETTE, PT_ULP_BIAS = 1, 2


def sample_disney(material, n, v, eta, seed, count):
    """count draws of the integrator's sample_disney: (count, 8) = direction xyz, pdf, F rgb, transmission"""
    h = lib()
    h.orc_sample_disney_n.restype = None
    h.orc_sample_disney_n.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    m = np.ascontiguousarray(material).view(np.float32)
    n3, v3 = np.asarray(n, np.float32), np.asarray(v, np.float32)
    out = np.zeros((count, 8), np.float32)
    h.orc_sample_disney_n(m.ctypes.data, n3.ctypes.data, v3.ctypes.data, eta, seed, count, out.ctypes.data)
    return out


def eval_disney(material, n, v, eta, dirs):
    """the integrator's eval_disney at each direction of dirs (m, 3): (m, 4) = F rgb, pdf"""
    h = lib()
    h.orc_eval_disney_n.restype = None
    h.orc_eval_disney_n.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                    ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    m = np.ascontiguousarray(material).view(np.float32)
    n3, v3 = np.asarray(n, np.float32), np.asarray(v, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    out = np.zeros((len(d), 4), np.float32)
    h.orc_eval_disney_n(m.ctypes.data, n3.ctypes.data, v3.ctypes.data, eta, d.ctypes.data, len(d), out.ctypes.data)
    return out


def texture_sample(rgba, uv, lam, sampler=(0, 0, 0, 0)):
    """SampleLevel of an (H, W, 4) uint8 image through its mip chain at uv (n, 2) and levels lam (n): (n, 4);
    sampler = (address mode u, address mode v, min filter, mag filter)"""
    h = lib()
    h.orc_texture_sample.restype = None
    h.orc_texture_sample.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_void_p]
    img = np.ascontiguousarray(rgba, np.uint8)
    t = _Texture(img.shape[1], img.shape[0], img.ctypes.data, *sampler)
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float32), (len(uv),)))
    out = np.zeros((len(uv), 4), np.float32)
    h.orc_texture_sample(ctypes.byref(t), uv.ctypes.data, lam.ctypes.data, len(uv), out.ctypes.data)
    return out


def pt_render(scene, camera, width, height, num_samples, max_depth, seed, max_path_beta=0.0, disable_material=0,
              nthreads=None, lights=None, mis_mode=0, suppress_refractive_caustics=0, test_flags=0, instanced=None):
    """one batch of the CPU reference integrator: (H, W, 8) float32 (rgb, coverage, depth, -, object bits, -);
    instanced = (mesh_first, PT_INSTANCE_DTYPE records) renders the two-level scene"""
    import vierkant_amd as vk
    from vierkant_amd import workloads as W
    h = lib()
    h.orc_pt_render.restype = None
    h.orc_pt_render.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    sc, keep = _pt_scene(scene, instanced)
    eye, fw, rt, up, thx, thy = W.camera_basis(camera, width, height)
    # the tracer's gather: positive intensity, no Area lights, order kept
    la = np.zeros(0, vk.LIGHT_DTYPE) if lights is None else np.asarray(lights, vk.LIGHT_DTYPE).reshape(-1)
    la = np.ascontiguousarray(la[(la["intensity"] > 0) & (la["type"] != vk.LIGHT_AREA)])
    pp = _PtParams(width, height, num_samples, max_depth, seed, disable_material, max_path_beta,
                   (ctypes.c_float * 3)(*eye), (ctypes.c_float * 3)(*fw), (ctypes.c_float * 3)(*rt),
                   (ctypes.c_float * 3)(*up), thx, thy, float(camera.aperture),
                   float(camera.focal_distance) if camera.focal_distance > 0 else 1.0,
                   la.ctypes.data if len(la) else None, len(la), mis_mode, suppress_refractive_caustics,
                   float(camera.fov_y), test_flags)
    out = np.zeros((height, width, 8), np.float32)
    h.orc_pt_render(ctypes.byref(sc), ctypes.byref(pp), out.ctypes.data, nthreads or threads())
    del keep
    return out
