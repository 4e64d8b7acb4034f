"""Writes the golden fixtures under tests/golden/ (run from the repository root; needs the oracle built with
`make -C oracle`, and oracle/_ref built from the reference checkout for the overlap vectors).

  tri_box_overlap.npz  4,000 triangle / AABB pairs and the 0/1 results of the reference's own
                       tri_box_overlap (src/triangle_intersection.h, compiled into oracle/_ref)
  tri_tri_overlap.npz  4,000 triangle pairs (1,000 of them coplanar, 500 sharing a vertex) and the results of
                       the reference's tri_tri_overlap_test_3d
  mt_icosphere3.npz    the jittered subdiv-3 icosphere, 2,048 rays and the oracle's closest hit (t/u/v bits,
                       prim) and any hit (finite tmax) records
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_binding as orc  # noqa: E402
from vierkant_amd import workloads as W  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def dyadic(rng, shape, lo, hi, bits=6):
    """random values on a 2^-bits grid, so touching configurations are exact"""
    return (np.round(rng.uniform(lo, hi, shape) * 2 ** bits) / 2 ** bits).astype(np.float32)


def tri_box_cases(rng, n=4000):
    tris = np.zeros((n, 9), np.float32)
    boxes = np.zeros((n, 6), np.float32)
    for i in range(n):
        kind = i % 4
        lo = dyadic(rng, 3, -1, 1)
        hi = lo + dyadic(rng, 3, 0, 1.5)
        if kind == 0:                          # random triangle near the box
            t = dyadic(rng, 9, -2, 2)
        elif kind == 1:                        # one vertex exactly on a box face
            t = dyadic(rng, 9, -2, 2)
            a = rng.integers(0, 3)
            t[a] = hi[a] if rng.uniform() < 0.5 else lo[a]
        elif kind == 2:                        # small triangle around a box corner
            c = np.where(rng.uniform(size=3) < 0.5, lo, hi)
            t = (np.tile(c, 3) + dyadic(rng, 9, -0.25, 0.25)).astype(np.float32)
        else:                                  # flat box or degenerate triangle
            t = dyadic(rng, 9, -2, 2)
            if rng.uniform() < 0.5:
                hi[rng.integers(0, 3)] = lo[rng.integers(0, 3)]
                hi = np.maximum(hi, lo)
            else:
                t[6:9] = t[0:3]
        tris[i], boxes[i, :3], boxes[i, 3:] = t, lo, hi
    return tris, boxes


def tri_tri_cases(rng, n=4000):
    a = dyadic(rng, (n, 9), -1, 1)
    b = dyadic(rng, (n, 9), -1, 1)
    # coplanar pairs: both in the plane z = c (or y = c)
    for i in range(0, 1000):
        c = dyadic(rng, 1, -1, 1)[0]
        ax = 2 if i % 2 else 1
        a[i, ax::3] = c
        b[i, ax::3] = c
    # pairs sharing a vertex
    for i in range(1000, 1500):
        b[i, 0:3] = a[i, 3:6]
    return a, b


def main():
    rng = np.random.default_rng(2024)
    ref = orc.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libvkref.so missing: build it from the reference checkout (make -C oracle)")
    tris, boxes = tri_box_cases(rng)
    tb = np.array([ref.ref_tri_box_overlap(tris[i].ctypes.data, boxes[i].ctypes.data) for i in range(len(tris))],
                  np.uint32)
    np.savez_compressed(os.path.join(OUT, "tri_box_overlap.npz"), tris=tris, boxes=boxes, expected=tb)
    a, b = tri_tri_cases(rng)
    tt = np.array([ref.ref_tri_tri_overlap(a[i].ctypes.data, b[i].ctypes.data) for i in range(len(a))], np.uint32)
    np.savez_compressed(os.path.join(OUT, "tri_tri_overlap.npz"), a=a, b=b, expected=tt)

    pos, idx = W.jittered_icosphere(3)
    flat = W.flat_triangles(pos, idx)
    o = rng.uniform(-2, 2, (2048, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(2048, 3)).astype(np.float32))
    rays = W.make_rays(o, d)
    closest = orc.brute_closest(flat, rays, 4)
    r2 = rays.copy()
    r2["tmax"] = rng.uniform(0, 3, len(r2)).astype(np.float32)
    anyhit = orc.brute_any(flat, r2, 4)
    np.savez_compressed(os.path.join(OUT, "mt_icosphere3.npz"), positions=pos, indices=idx, rays=rays,
                        closest=closest, rays_any=r2, occluded=anyhit)
    print(f"tri_box: {tb.sum()} of {len(tb)} overlap; tri_tri: {tt.sum()} of {len(tt)}; "
          f"closest hits: {(closest['prim'] != 0xFFFFFFFF).sum()} of 2048; occluded: {anyhit.sum()}")


if __name__ == "__main__":
    main()
