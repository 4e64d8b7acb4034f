"""GPU parity of the batched intersection tests and the brute-force loop against the CPU oracle.

Bit-exact: every INTERSECT/REJECT decision, every t/u/v bit pattern, every closest primitive index.
"""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from parity import assert_hits_equal, assert_results_equal, assert_tri_results_equal, bits
from vierkant_amd import workloads as W

pytestmark = pytest.mark.gpu


def _edge_rays_and_tris(rng, n):
    """random triangle/ray pairs, half aimed at the triangle, plus special values"""
    tris = rng.uniform(-2, 2, (n, 9)).astype(np.float32)
    v0 = tris[:, :3].astype(np.float64)
    bary = rng.uniform(-0.1, 1.1, (n, 2))
    p = v0 + bary[:, :1] * (tris[:, 3:6] - v0) + bary[:, 1:] * (tris[:, 6:9] - v0)
    o = rng.uniform(-4, 4, (n, 3))
    d = np.where(rng.uniform(size=(n, 1)) < 0.5, p - o, rng.normal(size=(n, 3)))
    d = W.glm_normalize(d.astype(np.float32))
    o = o.astype(np.float32)
    # special cases in the first rows
    k = 0
    specials = [
        ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0.25, 0.25, -1], [0, 0, 1]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0.5, 0.5, -1], [0, 0, 1]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0.0, 0.5, -1], [0, 0, 1]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0.25, 0.25, 1], [0, 0, 1]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0.25, 0.25, 0], [1, 0, 0]),
        ([0, 0, 0, 3e-5, 0, 0, 0, 3e-5, 0], [1e-5, 1e-5, -1], [0, 0, 1]),
        ([0, 0, 0, 4e-5, 0, 0, 0, 4e-5, 0], [1e-5, 1e-5, -1], [0, 0, 1]),
        ([0, 0, 0, 1e-38, 0, 0, 0, 1e-38, 0], [0, 0, -1], [0, 0, 1]),       # subnormal products
        ([0, 0, 0, 1, 0, 0, 0, 1, 0], [np.nan, 0, -1], [0, 0, 1]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0.25, 0.25, -1], [np.inf, 0, 1]),
        ([0, 0, 0, np.inf, 0, 0, 0, 1, 0], [0.25, 0.25, -1], [0, 0, 1]),
        ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0.25, 0.25, -1], [0, 0, 0]),         # zero direction
        ([1, 1, 1, 1, 1, 1, 1, 1, 1], [0, 0, 0], [0.5, 0.5, 0.5]),          # degenerate triangle
    ]
    for t9, oo, dd in specials:
        tris[k], o[k], d[k] = t9, oo, dd
        k += 1
    return tris, W.make_rays(o, d)


def test_pairwise_triangles_bit_exact(ctx):
    rng = np.random.default_rng(10)
    tris, rays = _edge_rays_and_tris(rng, 20000)
    got = ctx.intersect_triangles(tris, rays)
    want = np.zeros(len(rays), dtype=vk.TRI_RESULT_DTYPE)
    o, d = orc.ray_od(rays)
    for i in range(len(rays)):
        want[i] = orc.intersect_tri(tris[i], o[i], d[i])
    assert np.sum(want["type"] == 1) > 2000
    assert_tri_results_equal(got, want)


def test_pairwise_volumes_bit_exact(ctx):
    rng = np.random.default_rng(11)
    n = 8000
    o = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(n, 3)).astype(np.float32))
    d[:50, 1:] = 0.0           # axis-parallel rays exercise the |f| <= 1e-10 branch
    d[:50, 0] = 1.0
    rays = W.make_rays(o, d)
    lo = rng.uniform(-2, 1, (n, 3)).astype(np.float32)
    boxes = np.concatenate([lo, lo + rng.uniform(0, 2, (n, 3)).astype(np.float32)], axis=1)
    boxes[:20, 3:] = boxes[:20, :3]                  # flat boxes
    spheres = np.concatenate([rng.uniform(-2, 2, (n, 3)), rng.uniform(0, 1.5, (n, 1))], 1).astype(np.float32)
    nrm = W.glm_normalize(rng.normal(size=(n, 3)).astype(np.float32))
    planes = np.concatenate([nrm, rng.uniform(-2, 2, (n, 1)).astype(np.float32)], 1)
    axes = np.stack([np.linalg.qr(rng.normal(size=(3, 3)))[0] for _ in range(n)]).astype(np.float32)
    obbs = np.concatenate([rng.uniform(-2, 2, (n, 3)).astype(np.float32), axes.transpose(0, 2, 1).reshape(n, 9),
                           rng.uniform(0.05, 1.5, (n, 3)).astype(np.float32)], 1)
    for fn, ofn, prims in [(ctx.intersect_aabbs, orc.intersect_aabb, boxes),
                           (ctx.intersect_spheres, orc.intersect_sphere, spheres),
                           (ctx.intersect_planes, orc.intersect_plane, planes),
                           (ctx.intersect_obbs, orc.intersect_obb, obbs)]:
        prims = np.ascontiguousarray(prims)
        got = fn(prims, rays)
        want = np.zeros(n, dtype=vk.RESULT_DTYPE)
        for i in range(n):
            want[i] = ofn(prims[i], o[i], d[i])
        assert 100 < np.sum(want["type"] == 1) < n
        assert_results_equal(got, want)


def test_compute_aabb_bit_exact(ctx):
    rng = np.random.default_rng(12)
    for n in [1, 7, 1000, 300000]:
        pts = rng.normal(size=(n, 3)).astype(np.float32)
        assert bits(ctx.compute_aabb(pts)).tolist() == bits(orc.compute_aabb(pts)).tolist()
    # NaNs never win; signed zeros keep the first occurrence; infinities stay out of the initial box
    pts = np.array([[np.nan, 0.0, 5.0], [-0.0, -0.0, np.inf], [0.0, 3.0, -np.inf], [2.0, np.nan, 1.0]],
                   dtype=np.float32)
    assert bits(ctx.compute_aabb(pts)).tolist() == bits(orc.compute_aabb(pts)).tolist()
    assert bits(ctx.compute_aabb(np.zeros((0, 3), np.float32))).tolist() == \
        bits(orc.compute_aabb(np.zeros((0, 3), np.float32))).tolist()


@pytest.fixture(scope="module")
def icosphere_case():
    pos, idx, rays = W.icosphere_rays()
    tris = W.flat_triangles(pos, idx)
    return tris, rays, orc.brute_closest(tris, rays)


def test_brute_closest_icosphere_full_bit_exact(ctx, icosphere_case):
    """65,536 rays x 20,480 triangles: every record identical to the oracle"""
    tris, rays, want = icosphere_case
    got = ctx.brute_closest(tris, rays)
    assert np.sum(want["prim"] != 0xFFFFFFFF) > 3000   # about 6% of the random rays hit the sphere
    assert_hits_equal(got, want, "icosphere_case brute")


def test_brute_any_matches_oracle(ctx, icosphere_case):
    tris, rays, _ = icosphere_case
    r = rays[:8192].copy()
    r["tmax"] = np.random.default_rng(13).uniform(0.0, 6.0, len(r)).astype(np.float32)
    r["tmin"] = np.random.default_rng(14).uniform(-1.0, 1.0, len(r)).astype(np.float32)
    got = ctx.brute_any(tris, r)
    want = orc.brute_any(tris, r)
    assert 100 < want.sum() < len(r)
    assert np.array_equal(got, want)


def test_brute_closest_ties_resolve_to_lowest_index(ctx):
    rng = np.random.default_rng(15)
    base = rng.uniform(-1, 1, (200, 9)).astype(np.float32)
    tris = np.concatenate([base, base, base[::-1]])       # every triangle three times
    o = rng.uniform(-3, 3, (4000, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(4000, 3)).astype(np.float32))
    rays = W.make_rays(o, d)
    want = orc.brute_closest(tris, rays)
    got = ctx.brute_closest(tris, rays)
    assert np.sum(want["prim"] != 0xFFFFFFFF) > 100
    assert np.all(want["prim"][want["prim"] != 0xFFFFFFFF] < 200)
    assert_hits_equal(got, want, "ties")


def test_brute_adjacent_duplicates_and_remainders(ctx):
    """each triangle twice in a row, so ties fall inside one four-record loop step, and 602 triangles, so the
    triangle chunks leave records after the last full step: the lower index wins; any-hit agrees"""
    rng = np.random.default_rng(21)
    tris = np.repeat(rng.uniform(-1, 1, (301, 9)).astype(np.float32), 2, axis=0)
    o = rng.uniform(-3, 3, (3000, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(3000, 3)).astype(np.float32))
    rays = W.make_rays(o, d)
    want = orc.brute_closest(tris, rays)
    hit = want["prim"] != 0xFFFFFFFF
    assert hit.sum() > 100 and np.all(want["prim"][hit] % 2 == 0)
    assert_hits_equal(ctx.brute_closest(tris, rays), want, "adjacent duplicates")
    assert np.array_equal(ctx.brute_any(tris, rays), orc.brute_any(tris, rays))


@pytest.mark.parametrize("n_tri,n_ray", [(0, 10), (5, 0), (1, 1), (3, 300), (257, 1000)])
def test_brute_sizes(ctx, n_tri, n_ray):
    rng = np.random.default_rng(n_tri * 7 + n_ray)
    tris = rng.uniform(-1, 1, (n_tri, 9)).astype(np.float32)
    o = rng.uniform(-2, 2, (n_ray, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(n_ray, 3)).astype(np.float32)) if n_ray else np.zeros((0, 3), np.float32)
    rays = W.make_rays(o, d)
    got = ctx.brute_closest(tris, rays)
    want = orc.brute_closest(tris, rays) if n_ray else np.zeros(0, vk.HIT_DTYPE)
    assert_hits_equal(got, want, "sizes")
    assert np.array_equal(ctx.brute_any(tris, rays), orc.brute_any(tris, rays) if n_ray else np.zeros(0, np.uint8))


def test_device_pointer_path(ctx, icosphere_case):
    """device-resident inputs and outputs (VKRT_DEVICE_PTRS) give the same records"""
    tris, rays, want = icosphere_case
    dt = vk.DeviceArray.from_host(ctx, tris)
    dr = vk.DeviceArray.from_host(ctx, rays)
    out = vk.DeviceArray(ctx, len(rays), vk.HIT_DTYPE)
    ctx.brute_closest(dt, dr, out=out)
    assert_hits_equal(out.host(), want, "device pointers")
    occ = vk.DeviceArray(ctx, len(rays), np.uint8)
    ctx.brute_any(dt, dr, out=occ, flags=vk.ASYNC)
    ctx.sync()
    assert np.array_equal(occ.host(), (want["prim"] != 0xFFFFFFFF).astype(np.uint8))


@pytest.mark.parametrize("name", ["tri_box_overlap", "tri_tri_overlap"])
def test_overlap_kernels_match_reference_vectors(ctx, name):
    """the reference's own triangle_intersection.h results (golden vectors from oracle/_ref)"""
    import pathlib
    g = dict(np.load(pathlib.Path(__file__).resolve().parent / "golden" / f"{name}.npz"))
    if name == "tri_box_overlap":
        got = ctx.tri_box_overlap(g["tris"], g["boxes"])
    else:
        got = ctx.tri_tri_overlap(g["a"], g["b"])
    assert np.array_equal(got, g["expected"])


def test_overlap_kernels_match_reference_build_random(ctx):
    """fresh random pairs against the reference-compiled functions, when oracle/_ref is present"""
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(33)
    n = 20000
    a = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    b = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    b[: n // 4, 2::3] = a[: n // 4, 2:3]          # coplanar in z
    lo = rng.uniform(-1, 0.5, (n, 3)).astype(np.float32)
    boxes = np.concatenate([lo, lo + rng.uniform(0, 1, (n, 3)).astype(np.float32)], 1)
    want_tt = np.array([ref.ref_tri_tri_overlap(a[i].ctypes.data, b[i].ctypes.data) for i in range(n)], np.uint32)
    want_tb = np.array([ref.ref_tri_box_overlap(a[i].ctypes.data, boxes[i].ctypes.data) for i in range(n)], np.uint32)
    assert 0 < want_tt.sum() < n and 0 < want_tb.sum() < n
    assert np.array_equal(ctx.tri_tri_overlap(a, b), want_tt)
    assert np.array_equal(ctx.tri_box_overlap(a, boxes), want_tb)


@pytest.mark.parametrize("n", [1, 2, 3, 1000, 1025, 300000])
def test_bounding_sphere_bit_exact(ctx, n):
    """compute_bounding_sphere: the GPU's extremum choice and in-order growth equal the sequential oracle
    bit for bit; coordinates are quantized so that ties between extremum candidates are common"""
    rng = np.random.default_rng(n)
    pts = (np.round(rng.normal(size=(n, 3)) * rng.uniform(0.5, 5.0, size=3) * 8) / 8).astype(np.float32)
    got = ctx.compute_bounding_sphere(pts)
    want = orc.bounding_sphere(pts)
    assert got.view(np.uint32).tolist() == want.view(np.uint32).tolist()


def test_bounding_sphere_special_inputs(ctx):
    assert ctx.compute_bounding_sphere(np.zeros((0, 3), np.float32)).tolist() == [0.0, 0.0, 0.0, 0.0]
    pts = np.array([[0, 0, 0], [np.nan, 0, 0], [4, 0, 0]], np.float32)
    assert ctx.compute_bounding_sphere(pts).tolist() == orc.bounding_sphere(pts).tolist()
    pts = np.array([[np.nan, 0, 0], [1, 0, 0], [5, 0, 0]], np.float32)
    got, want = ctx.compute_bounding_sphere(pts), orc.bounding_sphere(pts)
    assert np.isnan(got[0]) and got[1:].tolist() == want[1:].tolist()


def test_brute_closest_adversarial_pairs_match_oracle(ctx):
    """the device's branch-joined triangle test takes the oracle's decisions and bits on the adversarial pairs
    (edge and vertex hits, det near 1e-9, huge and tiny triangles, non-finite inputs): one triangle per call,
    the ray's interval wide open so that every finite INTERSECT counts"""
    from mt_cases import adversarial_mt_cases
    checked = hits = 0
    for tri, o, d in adversarial_mt_cases():
        want = orc.intersect_tri(tri, o, d)
        ray = np.zeros(1, vk.RAY_DTYPE)
        ray["ox"], ray["oy"], ray["oz"] = o
        ray["dx"], ray["dy"], ray["dz"] = d
        ray["tmin"], ray["tmax"] = -3.4e38, 3.4e38
        got = ctx.brute_closest(tri[None, :], ray)[0]
        t = np.float32(want[1]) if want[0] else np.float32(np.nan)
        if want[0] and not (-3.4e38 < t < 3.4e38):
            continue                               # outside any ray interval: not a closest-hit candidate
        checked += 1
        assert (got["prim"] == 0) == bool(want[0]), (tri, o, d)
        if want[0]:
            hits += 1
            assert bits(np.array([got["t"], got["u"], got["v"]], np.float32)).tolist() == \
                bits(np.array(want[1:], np.float32)).tolist()
    assert checked > 600 and hits > 100
