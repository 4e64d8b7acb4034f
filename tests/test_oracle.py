"""Pins the CPU oracle (oracle/vkoracle.c) before results are compared with it.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a swapped cross-product
operand, a wrong sign or index) fails it: exact known answers computed by hand for representable inputs,
geometric closed forms, an FP64 twin, and independent implementations (the `xxhash` package, a closed-form
LCG jump, brute force in FP64).
"""
import numpy as np
import pytest

import oracle_binding as orc
from vierkant_amd import workloads as W

TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], dtype=np.float32)


# -------------------------------------------------------------------------------------------------
# ray - triangle (src/intersection.cpp:114-129)
# -------------------------------------------------------------------------------------------------
def test_mt_exact_known_answer():
    # e1=(1,0,0), e2=(0,1,0), p=cross(d,e2)=(-1,0,0), det=-1, u=0.25, q=(0,-1,-0.25), v=0.25, t=1
    assert orc.intersect_tri(TRI, [0.25, 0.25, -1], [0, 0, 1]) == (1, 1.0, 0.25, 0.25)


def test_mt_oblique_known_answer():
    # non-axis ray and a triangle where every cross-product term is non-zero: tri (1,0,0),(0,2,0),(0,0,4)
    # lies in the plane x + y/2 + z/4 = 1; the ray from the origin along (1, 1, 1) (unnormalized)
    tri = np.array([1, 0, 0, 0, 2, 0, 0, 0, 4], dtype=np.float32)
    r, t, u, v = orc.intersect_tri(tri, [0, 0, 0], [1, 1, 1])
    # exact: point p = t (1,1,1) with t (1 + 1/2 + 1/4) = 1 -> t = 4/7; p = v0 + u e1 + v e2 with
    # e1 = (-1,2,0), e2 = (-1,0,4): y = 2u -> u = 2/7, z = 4v -> v = 1/7
    assert r == 1
    assert t == pytest.approx(4 / 7, rel=2e-7)
    assert u == pytest.approx(2 / 7, rel=2e-7)
    assert v == pytest.approx(1 / 7, rel=2e-7)


@pytest.mark.parametrize("o,expect_uv", [
    ([0.0, 0.5, -1.0], (0.0, 0.5)),     # u = 0 edge (inclusive)
    ([0.5, 0.0, -1.0], (0.5, 0.0)),     # v = 0 edge (inclusive)
    ([0.5, 0.5, -1.0], (0.5, 0.5)),     # u + v = 1 edge (inclusive)
    ([0.0, 0.0, -1.0], (0.0, 0.0)),     # vertex v0
])
def test_mt_edges_inclusive(o, expect_uv):
    r, t, u, v = orc.intersect_tri(TRI, o, [0, 0, 1])
    assert r == 1 and t == 1.0 and (u, v) == expect_uv


@pytest.mark.parametrize("o", [[-2.0 ** -20, 0.5, -1.0], [0.5, -2.0 ** -20, -1.0], [0.5 + 2.0 ** -20, 0.5, -1.0]])
def test_mt_just_outside_rejects(o):
    assert orc.intersect_tri(TRI, o, [0, 0, 1])[0] == 0


def test_mt_two_sided_and_negative_t():
    flipped = np.array([0, 0, 0, 0, 1, 0, 1, 0, 0], dtype=np.float32)
    r, t, u, v = orc.intersect_tri(flipped, [0.25, 0.25, -1], [0, 0, 1])
    assert r == 1 and t == 1.0 and (u, v) == (0.25, 0.25)     # u, v swap roles with the winding
    # triangle behind the origin: INTERSECT with negative t (the reference has no t check)
    assert orc.intersect_tri(TRI, [0.25, 0.25, 1], [0, 0, 1]) == (1, -1.0, 0.25, 0.25)


def test_mt_parallel_and_det_threshold():
    assert orc.intersect_tri(TRI, [0.25, 0.25, 0], [1, 0, 0])[0] == 0   # det == 0
    small = np.array([0, 0, 0, 3e-5, 0, 0, 0, 3e-5, 0], dtype=np.float32)   # det = -9e-10 > -1e-9
    assert orc.intersect_tri(small, [1e-5, 1e-5, -1], [0, 0, 1])[0] == 0
    big = np.array([0, 0, 0, 4e-5, 0, 0, 0, 4e-5, 0], dtype=np.float32)     # det = -1.6e-9
    assert orc.intersect_tri(big, [1e-5, 1e-5, -1], [0, 0, 1])[0] == 1


def test_mt_barycentric_roundtrip_vs_fp64():
    """random triangle, random interior point with known (u, v), ray through it: the FP64 twin returns
    (u, v, t) to 1e-9 and the FP32 oracle to 1e-5 relative"""
    rng = np.random.default_rng(1)
    for _ in range(300):
        tri = rng.uniform(-2, 2, 9).astype(np.float32)
        v0, v1, v2 = tri[:3].astype(np.float64), tri[3:6].astype(np.float64), tri[6:].astype(np.float64)
        n = np.cross(v1 - v0, v2 - v0)
        if np.linalg.norm(n) < 0.5:
            continue
        u, v = rng.uniform(0.05, 0.45, 2)
        p = v0 + u * (v1 - v0) + v * (v2 - v0)
        o = (p + n / np.linalg.norm(n) * rng.uniform(0.5, 3) + rng.normal(0, 0.3, 3)).astype(np.float32)
        d = (p - o.astype(np.float64))
        d /= np.linalg.norm(d)
        d32 = d.astype(np.float32)
        r64, t64, u64, v64 = orc.intersect_tri_f64(tri, o, d32)
        assert r64 == 1
        assert u64 == pytest.approx(u, abs=1e-6) and v64 == pytest.approx(v, abs=1e-6)
        assert t64 == pytest.approx(np.linalg.norm(p - o.astype(np.float64)), rel=1e-6)
        r32, t32, u32, v32 = orc.intersect_tri(tri, o, d32)
        assert r32 == 1
        assert t32 == pytest.approx(t64, rel=1e-5) and u32 == pytest.approx(u64, rel=1e-4, abs=1e-6)
        assert v32 == pytest.approx(v64, rel=1e-4, abs=1e-6)


def test_brute_force_matches_fp64_on_separated_triangles():
    """small scenes of well separated triangles: the FP32 oracle's closest triangle equals the FP64 one"""
    import vierkant_amd.workloads as W
    rng = np.random.default_rng(2)
    tris = []
    for k in range(40):
        c = np.array([k % 5, (k // 5) % 4, k // 20], dtype=np.float64) * 2.0
        tris.append((c + rng.uniform(-0.8, 0.8, (3, 3))).reshape(9))
    tris = np.asarray(tris, dtype=np.float32)
    o = rng.uniform(-2, 10, (256, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(256, 3)).astype(np.float32))
    rays = W.make_rays(o, d)
    hits = orc.brute_closest(tris, rays, 2)
    for i in range(len(rays)):
        prim, t, u, v = orc.closest_f64(tris, rays[i])
        assert hits["prim"][i] == prim
        if prim != 0xFFFFFFFF:
            assert hits["t"][i] == pytest.approx(t, rel=1e-5)


# -------------------------------------------------------------------------------------------------
# ray - AABB / OBB / sphere / plane (src/intersection.cpp:152-184, 308-312, 133-148, 10-21)
# -------------------------------------------------------------------------------------------------
UNIT = np.array([0, 0, 0, 1, 1, 1], dtype=np.float32)


@pytest.mark.parametrize("o,d,expect", [
    ([-1, 0.5, 0.5], [1, 0, 0], (1, 1.0)),            # outside -> entry distance
    ([0.5, 0.5, 0.5], [1, 0, 0], (1, 0.5)),           # inside -> exit distance
    ([0.5, 0.5, 0.5], [0, 0, -1], (1, 0.5)),
    ([-1, 2, 0.5], [1, 0, 0], (0, 0.0)),              # parallel, outside the y slab
    ([-1, 1, 0.5], [1, 0, 0], (1, 1.0)),              # parallel on the face: inclusive
    ([2, 0.5, 0.5], [1, 0, 0], (0, 0.0)),             # box behind
    ([-1, -1, 0.5], [0.6, 0.8, 0], (1, 5.0 / 3.0)),   # oblique: entry = max(1/0.6, 1/0.8)
])
def test_aabb_closed_form(o, d, expect):
    r, dist = orc.intersect_aabb(UNIT, o, d)
    assert (r, dist) == (expect[0], pytest.approx(expect[1], rel=1e-7))


def test_obb_rotated_closed_form():
    # a box rotated 90 degrees about z (axes permuted) behaves like the AABB with swapped half lengths
    obb = np.array([0, 0, 5, 0, 1, 0, -1, 0, 0, 0, 0, 1, 2, 1, 1], dtype=np.float32)
    # half length 2 along world y, 1 along world x
    assert orc.intersect_obb(obb, [0, -10, 5], [0, 1, 0]) == (1, 8.0)
    assert orc.intersect_obb(obb, [-10, 0, 5], [1, 0, 0]) == (1, 9.0)


@pytest.mark.parametrize("o,d,expect", [
    ([0, 0, 0], [0, 0, 1], (1, 4.0)),      # front hit
    ([0, 0, 5], [0, 0, 1], (1, 1.0)),      # origin inside -> far hit
    ([0, 0, 10], [0, 0, 1], (0, 0.0)),     # sphere behind
    ([1, 0, 0], [0, 0, 1], (1, 5.0)),      # tangent
    ([1.5, 0, 0], [0, 0, 1], (0, 0.0)),    # miss
])
def test_sphere_closed_form(o, d, expect):
    assert orc.intersect_sphere([0, 0, 5, 1], o, d) == expect


@pytest.mark.parametrize("o,d,expect", [
    ([0, 2, 0], [0, -1, 0], (1, 2.0)),
    ([0, 2, 0], [0, 1, 0], (0, 0.0)),      # negative distance rejected
    ([0, 2, 0], [1, 0, 0], (0, 0.0)),      # parallel (|denom| <= 1e-6)
    ([3, 2, 0], [0.6, -0.8, 0], (1, 2.5)),
])
def test_plane_closed_form(o, d, expect):
    assert orc.intersect_plane([0, 1, 0, 0], o, d) == (expect[0], pytest.approx(expect[1], rel=1e-7))


def test_compute_aabb_matches_numpy_and_empty():
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(1000, 3)).astype(np.float32)
    out = orc.compute_aabb(pts)
    assert np.array_equal(out[:3], pts.min(0)) and np.array_equal(out[3:], pts.max(0))
    empty = orc.compute_aabb(np.zeros((0, 3), np.float32))
    f = np.finfo(np.float32).max
    assert np.array_equal(empty, [f, f, f, -f, -f, -f])


# -------------------------------------------------------------------------------------------------
# procedural geometry (src/Geometry.cpp)
# -------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("subdiv", [0, 1, 2, 3])
def test_icosphere_closed_form(subdiv):
    pos, idx = orc.icosphere(1.0, subdiv)
    assert len(idx) == 3 * 20 * 4 ** subdiv
    assert len(pos) == 12 + 3 * sum(20 * 4 ** k for k in range(subdiv))
    r = np.linalg.norm(pos.astype(np.float64), axis=1)
    assert np.allclose(r, 1.0, atol=3e-7)
    # every triangle is small and outward-consistent: face normals point away from the center
    tri = pos[idx.reshape(-1, 3)].astype(np.float64)
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    c = tri.mean(1)
    assert np.all(np.sum(n * c, axis=1) != 0)
    s = np.sign(np.sum(n * c, axis=1))
    assert np.all(s == s[0])


def test_icosphere_subdivision_structure():
    """level-1 vertex k is radius * normalize(va + vc), k+1 normalize(va + vb), k+2 normalize(vb + vc)
    for the parent triangle (a, b, c); children follow {a,k+1,k, k+1,b,k+2, k,k+2,c, k,k+1,k+2}"""
    p0, i0 = orc.icosphere(2.0, 0)
    p1, i1 = orc.icosphere(2.0, 1)
    assert np.array_equal(p1[:12], p0)
    for j in range(20):
        a, b, c = i0[3 * j: 3 * j + 3]
        k = 12 + 3 * j
        for slot, (x, y) in zip(range(3), [(a, c), (a, b), (b, c)]):
            m = p0[x].astype(np.float64) + p0[y].astype(np.float64)
            assert np.allclose(p1[k + slot], 2.0 * m / np.linalg.norm(m), atol=1e-6)
        assert list(i1[12 * j: 12 * j + 12]) == [a, k + 1, k, k + 1, b, k + 2, k, k + 2, c, k, k + 1, k + 2]


def test_plane_layout():
    pos, idx = orc.plane(10.0, 10.0, 4, 2)
    assert pos.shape == (15, 3) and len(idx) == 4 * 2 * 6
    assert np.array_equal(pos[0], [-5, 0, -5]) and np.array_equal(pos[-1], [5, 0, 5])
    assert list(idx[:6]) == [0, 5, 6, 6, 1, 0]     # a=0 b=5 c=6 d=1: (a,b,c),(c,d,a)


# -------------------------------------------------------------------------------------------------
# RNG of the path tracer (shaders/slang/utils/{hash,random,sampling}.slang)
# -------------------------------------------------------------------------------------------------
def test_xxhash32_is_xxh32_of_four_bytes():
    """the shader's xxhash32(lhs, rhs) is XXH32 of the 4-byte little-endian rhs with seed lhs - 4 (the
    shader omits the input-length term); checked against the independent xxhash library"""
    xxhash = pytest.importorskip("xxhash")
    rng = np.random.default_rng(4)
    for a, b in [(0, 0), (1, 2), (0xFFFFFFFF, 7)] + [tuple(map(int, x)) for x in rng.integers(0, 2 ** 32, (200, 2))]:
        expect = xxhash.xxh32_intdigest(int(b).to_bytes(4, "little"), seed=(a - 4) % 2 ** 32)
        assert orc.xxhash32(a, b) == expect


def test_lcg_matches_closed_form_jump():
    a, c, m = 1664525, 1013904223, 2 ** 32
    seq, final = orc.rnd_sequence(12345, 1000)
    # x_n = a^n x_0 + c (a^n - 1) / (a - 1) mod 2^32, the geometric sum taken mod (a - 1) 2^32
    an = pow(a, 1000, m * (a - 1))
    xn = (pow(a, 1000, m) * 12345 + c * ((an - 1) // (a - 1))) % m
    assert final == xn
    assert all(0.0 <= x < 1.0 for x in seq)
    state = 12345
    first = (a * state + c) % m
    assert seq[0] == (first & 0xFFFFFF) / 2 ** 24


def test_hammersley_closed_form():
    for k in range(8):
        assert orc.hammersley(2 ** k, 1)[1] == 2.0 ** -(k + 1)
    assert tuple(orc.hammersley(3, 4)) == (0.75, 0.75)
    assert tuple(orc.hammersley(0, 4)) == (0.0, 0.0)


# ---- compute_bounding_sphere / compute_centroid (src/intersection.cpp:25-110) ------------------------------

def test_bounding_sphere_two_points_exact():
    """two points: the seed diameter is the segment, no point lies outside: center = midpoint, r = |p2-p1|/2"""
    s = orc.bounding_sphere(np.array([[1.0, 2.0, 3.0], [5.0, 2.0, 3.0]], np.float32))
    assert s.tolist() == [3.0, 2.0, 3.0, 2.0]


def test_bounding_sphere_hand_computed_growth():
    """x extremes (-2,0,0),(2,0,0) (squared extent 16 beats y's 13) seed c = 0, r = 2; (0,3,0) lies outside
    with d = 3: k = 0.5 + (2/3)/2, c = (0,3,0)(1 - k) + 0 k, r = (2 + 3)/2 = 2.5"""
    pts = np.array([[-2, 0, 0], [2, 0, 0], [0, 3, 0]], np.float32)
    s = orc.bounding_sphere(pts)
    k = np.float32(0.5) + (np.float32(2.0) / np.float32(3.0)) / np.float32(2.0)
    assert s[0] == 0.0 and s[2] == 0.0 and s[3] == np.float32(2.5)
    assert s[1] == np.float32(3.0) * (np.float32(1.0) - k)


def test_bounding_sphere_contains_all_points():
    rng = np.random.default_rng(8)
    for n in (1, 3, 17, 1000, 20000):
        pts = (rng.normal(size=(n, 3)) * rng.uniform(0.1, 10, size=3)).astype(np.float32)
        s = orc.bounding_sphere(pts)
        d = np.linalg.norm(pts.astype(np.float64) - s[:3].astype(np.float64), axis=1)
        assert (d <= s[3] * (1 + 1e-5) + 1e-6).all()
        lo, hi = pts.min(0), pts.max(0)
        assert s[3] <= np.linalg.norm(hi - lo)            # never above the box diagonal


def test_bounding_sphere_nan_never_wins():
    """a NaN coordinate fails every strict comparison: it is never an extremum and never grows the sphere"""
    pts = np.array([[0, 0, 0], [np.nan, 0, 0], [4, 0, 0]], np.float32)
    assert orc.bounding_sphere(pts).tolist() == [2.0, 0.0, 0.0, 2.0]
    # a NaN at index 0 stays the candidate of that axis; every seed distance is then NaN, so axis 0 with
    # squared extent 0 seeds c = (p0 + p0) / 2 and r = 0, and no NaN distance exceeds r^2
    s = orc.bounding_sphere(np.array([[np.nan, 0, 0], [1, 0, 0], [5, 0, 0]], np.float32))
    assert np.isnan(s[0]) and s[1:].tolist() == [0.0, 0.0, 0.0]


def test_centroid_sequential_mean():
    pts = np.array([[1, 2, 3], [3, 2, 1], [2, 2, 2]], np.float32)
    assert orc.centroid(pts).tolist() == [2.0, 2.0, 2.0]
    assert orc.centroid(np.zeros((0, 3), np.float32)).tolist() == [0.0, 0.0, 0.0]


# ---- two-level scene (instances with world-to-object transforms) ------------------------------------------

def _instances(mats, meshes, customs=None):
    from vierkant_amd import INSTANCE_DTYPE
    inst = np.zeros(len(mats), INSTANCE_DTYPE)
    for k, (o2w, w2o) in enumerate(mats):
        inst[k]["object_to_world"] = np.asarray(o2w, np.float32).reshape(12)
        inst[k]["world_to_object"] = np.asarray(w2o, np.float32).reshape(12)
        inst[k]["blas"] = meshes[k]
        inst[k]["custom_index"] = 100 + k if customs is None else customs[k]
    return inst


def _affine(scale=1.0, t=(0.0, 0.0, 0.0)):
    m = np.zeros((3, 4), np.float64)
    m[:, :3] = np.eye(3) * scale
    m[:, 3] = t
    w = np.zeros((3, 4), np.float64)
    w[:, :3] = np.eye(3) / scale
    w[:, 3] = -np.asarray(t) / scale
    return m, w


def _small_mesh(seed):
    rng = np.random.default_rng(seed)
    return (np.round(rng.uniform(-1, 1, (40, 9)) * 8) / 8).astype(np.float32)


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = (np.round(rng.uniform(-3, 3, (n, 3)) * 8) / 8).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    return W.make_rays(o, W.glm_normalize(d))


def test_scene_identity_instance_equals_brute_force():
    tris = _small_mesh(1)
    rays = _rays(500, 2)
    got = orc.scene_closest([tris], _instances([_affine()], [0]), rays)
    want = orc.brute_closest(tris.reshape(-1), rays)
    assert got["prim"].tolist() == want["prim"].tolist()
    assert got["t"].view(np.uint32).tolist() == want["t"].view(np.uint32).tolist()
    hit = want["prim"] != 0xFFFFFFFF
    assert (got["instance"][hit] == 0).all() and (got["instance"][~hit] == 0xFFFFFFFF).all()
    assert (got["custom_index"][hit] == 100).all()


def test_scene_ties_go_to_the_lower_instance():
    tris = _small_mesh(3)
    rays = _rays(300, 4)
    got = orc.scene_closest([tris], _instances([_affine(), _affine()], [0, 0]), rays)
    hit = got["prim"] != 0xFFFFFFFF
    assert hit.any() and (got["instance"][hit] == 0).all()


def test_scene_power_of_two_scale_and_grid_translation_are_exact():
    """object->world = 2 x + (1, -2, 0.5): world rays map to object rays exactly, so t, u, v equal the
    brute force over the world-space triangles (whose coordinates are exact too)"""
    tris = _small_mesh(5)
    m, w = _affine(2.0, (1.0, -2.0, 0.5))
    world = tris.reshape(-1, 3, 3) * 2 + np.array([1.0, -2.0, 0.5], np.float32)
    rays = _rays(600, 6)
    got = orc.scene_closest([tris], _instances([(m, w)], [0]), rays)
    want = orc.brute_closest(world.astype(np.float32).reshape(-1), rays)
    assert got["prim"].tolist() == want["prim"].tolist()
    assert got["t"].view(np.uint32).tolist() == want["t"].view(np.uint32).tolist()
    assert got["u"].view(np.uint32).tolist() == want["u"].view(np.uint32).tolist()


# ---- transmittance shadow rays (visibility_test.slang:94-165) ----------------------------------------------

def _quad_scene(zs, trans, color, two_sided=True):
    """unit quads in planes z = zs[k] (facing +z), each with its own material"""
    import vierkant_amd as vk
    pos, idx, tri_obj = [], [], []
    for k, z in enumerate(zs):
        base = len(pos)
        pos += [(-1, -1, z), (1, -1, z), (1, 1, z), (-1, 1, z)]
        idx += [base, base + 1, base + 2, base, base + 2, base + 3]
        tri_obj += [k, k]
    mats = np.array([vk.material(color, transmission=t, two_sided=two_sided) for t in trans], dtype=vk.MATERIAL_DTYPE)
    pos = np.array(pos, np.float32)
    return {"positions": pos, "normals": np.tile(np.array([0, 0, 1], np.float32), (len(pos), 1)),
            "indices": np.array(idx, np.uint32), "tri_object": np.array(tri_obj, np.uint32),
            "object_material": np.arange(len(zs), dtype=np.uint32), "materials": mats, "environment": 1.0}


def _down_rays(n=4, tmax=np.inf):
    o = np.tile(np.array([[0.25, 0.125, 5.0]], np.float32), (n, 1))
    d = np.tile(np.array([[0.0, 0.0, -1.0]], np.float32), (n, 1))
    return W.make_rays(o, d, tmin=0.0, tmax=tmax)


def test_transmittance_closed_forms():
    sc = _quad_scene([1.0, 2.0], [0.5, 0.5], (1.0, 0.5, 0.25))
    t = orc.transmittance(sc, _down_rays())
    assert np.array_equal(t[:, :3], np.tile([0.25, 0.0625, 0.015625], (4, 1)))
    # an opaque plate in the interval blocks; beyond tmax it does not count
    sc = _quad_scene([1.0, 2.0, -1.0], [0.5, 0.5, 0.0], (1.0, 0.5, 0.25))
    assert (orc.transmittance(sc, _down_rays())[:, :3] == 0).all()
    assert np.array_equal(orc.transmittance(sc, _down_rays(tmax=5.5))[:, :3], np.tile([0.25, 0.0625, 0.015625], (4, 1)))
    # nothing hit: fully visible
    assert (orc.transmittance(sc, W.make_rays(np.array([[5, 5, 5]], np.float32),
                                              np.array([[0, 0, -1]], np.float32)))[:, :3] == 1).all()


def test_transmittance_back_face_culling():
    """the plates face +z: a ray going +z meets their back faces, culled unless two-sided"""
    up = W.make_rays(np.array([[0.25, 0.125, -5.0]], np.float32), np.array([[0, 0, 1.0]], np.float32))
    one = _quad_scene([1.0, 2.0], [0.0, 0.5], (1.0, 1.0, 1.0), two_sided=False)
    assert (orc.transmittance(one, up)[:, :3] == 0).all()
    assert (orc.transmittance(one, up, cull=True)[:, :3] == 1).all()
    two = _quad_scene([1.0, 2.0], [0.0, 0.5], (1.0, 1.0, 1.0), two_sided=True)
    assert (orc.transmittance(two, up, cull=True)[:, :3] == 0).all()
