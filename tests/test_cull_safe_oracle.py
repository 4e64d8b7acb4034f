"""The oracle's culling criterion (oracle/vkoracle.c orc_cull_safe, orc_brute_closest_cull_safe) against
float64 slab geometry worked out here. The criterion restates the library's conservative box test on the
winning triangle's own padded box (vertex min/max grown by 2^-16 of each coordinate's magnitude): when that
box is entered with best t = the winner's t, every enclosing BVH box is entered as well, so any BVH the
library builds returns the brute-force record. The tests pin the two directions that make the criterion
useful: it never rejects a hit lying inside the exact interval of the padded box (conservative), and it
does reject a t clearly before the box, and a tmin clearly past it (not vacuous)."""
import numpy as np

import oracle_binding as orc
from vierkant_amd import workloads as W

PAD = 2.0 ** -16


def exact_interval(tri9, ray8, pad, box=None):
    """float64 slab interval of the triangle's vertex box grown by `pad` of each coordinate's magnitude (or
    of the given (lo, hi) box)"""
    v = np.asarray(tri9, np.float64).reshape(3, 3)
    o = np.asarray(ray8[0:3], np.float64)
    d = np.asarray(ray8[4:7], np.float64)
    mn, mx = v.min(axis=0), v.max(axis=0)
    lo, hi = (mn - np.abs(mn) * pad, mx + np.abs(mx) * pad) if box is None else box
    tn, tf = -np.inf, np.inf
    for a in range(3):
        if d[a] == 0.0:
            if not (lo[a] <= o[a] <= hi[a]):
                return np.inf, -np.inf
            continue
        t0, t1 = (lo[a] - o[a]) / d[a], (hi[a] - o[a]) / d[a]
        tn, tf = max(tn, min(t0, t1)), min(tf, max(t0, t1))
    return tn, tf


def ray8(o, d, tmin=0.0, tmax=np.inf):
    return np.array([*o, tmin, *d, tmax], np.float32)


def random_hits(n, seed):
    """rays from 2..5 units away aimed at a random interior point of a random unit-sized triangle"""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        tri = rng.uniform(-1.0, 1.0, 9).astype(np.float32) * np.float32(rng.uniform(0.01, 10.0))
        b = rng.dirichlet([1.0, 1.0, 1.0])
        target = (tri.reshape(3, 3).astype(np.float64) * b[:, None]).sum(axis=0)
        o = target + rng.normal(size=3) * rng.uniform(2.0, 5.0) * np.abs(tri).max()
        d = W.glm_normalize((target - o)[None, :].astype(np.float32))[0]
        r = ray8(o.astype(np.float32), d)
        hit, t, _, _ = orc.intersect_tri(tri, r[0:3], r[4:7])
        if hit == 1 and t > 0:
            out.append((tri, r, np.float32(t)))
    return out


def test_conservative_on_random_hits():
    """a hit t inside the exact interval of the padded box (pad 2^-16 less the fp32 rounding of the box
    corners) always passes; nearly every random hit is such a hit"""
    inside = 0
    hits = random_hits(3000, 1)
    for tri, r, t in hits:
        tn, tf = exact_interval(tri, r, PAD - 2.0 ** -23)
        if tn <= t <= tf:
            inside += 1
            assert orc.cull_safe(tri, r, t) == 1, (tri, r, t, tn, tf)
    assert inside >= 0.99 * len(hits)


def test_axis_planes_through_the_origin():
    """triangles in the plane z = 0 have boxes of zero thickness there (no pad on a zero coordinate): the
    per-ray offsets and the interval pad still admit the reference test's t, whatever its rounding"""
    rng = np.random.default_rng(2)
    tri = np.array([-1, -1, 0, 1, -1, 0, 0, 1, 0], np.float32)
    n_hit = 0
    for _ in range(2000):
        o = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(0.1, 7.0)], np.float32)
        target = np.array([rng.uniform(-0.4, 0.4), rng.uniform(-0.6, 0.4), 0.0])
        d = W.glm_normalize((target - o)[None, :].astype(np.float32))[0]
        r = ray8(o, d)
        hit, t, _, _ = orc.intersect_tri(tri, r[0:3], r[4:7])
        if hit == 1 and t > 0:
            n_hit += 1
            assert orc.cull_safe(tri, r, t) == 1, (o, d, t)
    assert n_hit > 1500


def plane_hits(z, elevation, n, seed):
    """rays onto the triangle (-1,-1,z) (1,-1,z) (0,1,z) at elevation angles drawn from `elevation`
    (radians above the plane), from 0.5..3 units away"""
    rng = np.random.default_rng(seed)
    tri = np.array([-1, -1, z, 1, -1, z, 0, 1, z], np.float32)
    out = []
    while len(out) < n:
        target = np.array([rng.uniform(-0.4, 0.4), rng.uniform(-0.6, 0.4), z])
        el, az, dist = rng.uniform(*elevation), rng.uniform(0, 2 * np.pi), rng.uniform(0.5, 3.0)
        back = np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)]) * dist
        o = (target + back).astype(np.float32)
        d = W.glm_normalize((target - o)[None, :].astype(np.float32))[0]
        r = ray8(o, d)
        hit, t, _, _ = orc.intersect_tri(tri, r[0:3], r[4:7])
        if hit == 1 and t > 0:
            out.append((tri, r, np.float32(t)))
    return out


def test_axis_planes_off_the_origin():
    """a triangle in the plane z = c has a box only 2 |c| 2^-16 thick in z: the hit sits between its walls
    and passes (a pad of the wrong sign leaves an empty box)"""
    for k, z in enumerate((0.75, -1.3, 1000.0)):
        for tri, r, t in plane_hits(z, (0.3, 1.5), 300, 10 + k):
            assert orc.cull_safe(tri, r, t) == 1, (z, r, t)


def test_shallow_rays_on_a_plane_through_the_origin():
    """rays 1 to 10 degrees above the plane z = 0 (a box of zero thickness in z): the reference test's t
    carries a relative error beyond the per-plane offsets' 2^-20 there, which the interval pad absorbs"""
    hits = plane_hits(0.0, (np.radians(1.0), np.radians(10.0)), 3000, 20)
    assert all(orc.cull_safe(tri, r, t) == 1 for tri, r, t in hits)


def box_triangle(lo, hi):
    """a triangle whose vertex box is [lo, hi]: the two corners and a point between them"""
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    return np.concatenate([lo, hi, (lo + (hi - lo) * np.float32(0.3)).astype(np.float32)]).astype(np.float32)


def boundary_cases(origin_scale, box_scale, n, seed, near=False):
    """unit-sized boxes centred within box_scale of the world origin; ray origins within origin_scale of the
    world origin, or (near) 2..4 units from the box"""
    rng = np.random.default_rng(seed)
    for _ in range(n):
        c = rng.uniform(-1, 1, 3) * box_scale
        lo = (c - rng.uniform(0.01, 1.0, 3)).astype(np.float32)
        hi = (c + rng.uniform(0.01, 1.0, 3)).astype(np.float32)
        tri = box_triangle(lo, hi)
        if near:
            u = rng.normal(size=3)
            o = (c + u / np.linalg.norm(u) * rng.uniform(2.0, 4.0)).astype(np.float32)
        else:
            o = (rng.uniform(-1, 1, 3) * origin_scale).astype(np.float32)
        d = W.glm_normalize((c + rng.uniform(-0.3, 0.3, 3) - o)[None, :].astype(np.float32))[0]
        yield tri, ray8(o, d)


def test_the_exact_padded_interval_ends_pass():
    """best t at the exact entry of the stored padded box, and tmin at its exact exit (float64, rounded to the
    nearest float), pass: the float interval contains the exact one. Rays from the world origin to boxes far
    away need the interval pad (no per-plane offset there); rays starting a few units from a box 3000 units
    out need the per-plane offsets (o inv then rounds by more than the pad's share of t)"""
    n = 0
    for origin_scale, box_scale, near in ((0.0, 50.0, False), (3000.0, 2.0, False), (10.0, 10.0, False),
                                          (0.0, 3000.0, True)):
        for tri, r in boundary_cases(origin_scale, box_scale, 1000, int(origin_scale + box_scale), near):
            # the stored box: the float corners grown by 2^-16 of their magnitude, rounded to float once
            v = tri.reshape(3, 3)
            mn, mx = v.min(axis=0), v.max(axis=0)
            pad = np.float32(PAD)
            box = ((mn - np.abs(mn) * pad).astype(np.float64), (mx + np.abs(mx) * pad).astype(np.float64))
            tn, tf = exact_interval(tri, r, PAD, box)
            if not (0.0 < tn < tf):
                continue
            n += 1
            assert orc.cull_safe(tri, r, np.float32(tn)) == 1, (tri, r, tn)
            late = r.copy()
            late[3] = np.float32(tf)
            assert orc.cull_safe(tri, late, np.float32(np.inf)) == 1, (tri, r, tf)
    assert n > 3500


def test_rejects_a_t_before_the_box_and_a_tmin_past_it():
    """best t clearly before the box's entry (below the exact entry of a box grown 16 times as much) culls
    it; so does a tmin clearly past the exit"""
    checked = 0
    for tri, r, t in random_hits(1000, 3):
        tn, tf = exact_interval(tri, r, 2.0 ** -12)
        if tn > 1e-3:
            checked += 1
            assert orc.cull_safe(tri, r, np.float32(tn * (1.0 - 2.0 ** -10))) == 0
            late = r.copy()
            late[3] = np.float32(tf * (1.0 + 2.0 ** -10))
            assert orc.cull_safe(tri, late, t) == 0
    assert checked > 900


def test_brute_force_flags_are_the_winners_criterion():
    """orc_brute_closest_cull_safe returns orc_brute_closest's records, and per ray the criterion of the
    winning triangle at its t (1 for a miss)"""
    pos, idx = W.jittered_icosphere(subdiv=2)
    tris = W.flat_triangles(pos, idx)
    rng = np.random.default_rng(4)
    o = rng.uniform(-3, 3, (512, 3)).astype(np.float32)
    d = W.glm_normalize((rng.normal(scale=0.6, size=(512, 3)) - o).astype(np.float32))
    rays = W.make_rays(o, d)
    hits, safe = orc.brute_closest_cull_safe(tris, rays)
    want = orc.brute_closest(tris, rays)
    assert np.array_equal(hits.view(np.uint8), want.view(np.uint8))
    flat = rays.view(np.float32).reshape(-1, 8)
    n_hit = 0
    for i in range(len(rays)):
        p = int(hits["prim"][i])
        if p == 0xFFFFFFFF:
            assert safe[i] == 1
        else:
            n_hit += 1
            assert safe[i] == orc.cull_safe(tris[p], flat[i], hits["t"][i])
    assert n_hit > 100 and safe.all()
