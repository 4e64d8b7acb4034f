"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py): the oracle reproduces its own
vectors bit for bit, the reference-compiled triangle_intersection.h reproduces the overlap vectors, and the
product's intersection header compiled for the host (tests/cpp/rt_math_host.cpp) gives the same results."""
import ctypes
import pathlib

import numpy as np
import pytest

import oracle_binding as orc
from mt_cases import adversarial_mt_cases

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
HOST = pathlib.Path(__file__).resolve().parent / "cpp" / "librt_math_host.so"


def host():
    if not HOST.exists():
        pytest.skip("tests/cpp/librt_math_host.so not built")
    h = ctypes.CDLL(str(HOST))
    h.host_tri_box_overlap.restype = ctypes.c_uint32
    h.host_tri_box_overlap.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    h.host_tri_tri_overlap.restype = ctypes.c_uint32
    h.host_tri_tri_overlap.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    h.host_mt.restype = ctypes.c_int
    h.host_mt.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p]
    return h


def test_oracle_reproduces_mt_fixture():
    g = dict(np.load(GOLD / "mt_icosphere3.npz"))
    flat = g["positions"][g["indices"].reshape(-1, 3)].reshape(-1, 9).astype(np.float32)
    got = orc.brute_closest(flat, g["rays"], 2)
    assert got.tobytes() == g["closest"].tobytes()
    assert np.array_equal(orc.brute_any(flat, g["rays_any"], 2), g["occluded"])


@pytest.mark.parametrize("name", ["tri_box_overlap", "tri_tri_overlap"])
def test_reference_build_reproduces_overlap_fixture(name):
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    g = dict(np.load(GOLD / f"{name}.npz"))
    if name == "tri_box_overlap":
        a, b = g["tris"], g["boxes"]
        fn = ref.ref_tri_box_overlap
    else:
        a, b = g["a"], g["b"]
        fn = ref.ref_tri_tri_overlap
    got = np.array([fn(np.ascontiguousarray(a[i]).ctypes.data, np.ascontiguousarray(b[i]).ctypes.data)
                    for i in range(len(a))], np.uint32)
    assert np.array_equal(got, g["expected"])


def test_host_header_overlaps_match_reference_vectors():
    h = host()
    g = dict(np.load(GOLD / "tri_box_overlap.npz"))     # load once: NpzFile re-reads on every access
    tris, boxes = g["tris"], g["boxes"]
    got = np.array([h.host_tri_box_overlap(tris[i].ctypes.data, boxes[i].ctypes.data) for i in range(len(tris))],
                   np.uint32)
    assert np.array_equal(got, g["expected"])
    g = dict(np.load(GOLD / "tri_tri_overlap.npz"))
    a, b = g["a"], g["b"]
    got = np.array([h.host_tri_tri_overlap(a[i].ctypes.data, b[i].ctypes.data) for i in range(len(a))], np.uint32)
    assert np.array_equal(got, g["expected"])


def test_host_header_mt_matches_oracle_bitwise():
    """mt_reference and the division-deferring mt_fast both equal the oracle's intersect(Triangle, Ray)"""
    h = host()
    g = dict(np.load(GOLD / "mt_icosphere3.npz"))
    flat = g["positions"][g["indices"].reshape(-1, 3)].reshape(-1, 9).astype(np.float32)
    rng = np.random.default_rng(5)
    rays = g["rays"]
    o, d = orc.ray_od(rays)
    tuv = np.zeros(3, np.float32)
    for k in range(3000):
        i, j = rng.integers(0, len(rays)), rng.integers(0, len(flat))
        if k < 300:      # rays that do hit their triangle
            j = int(g["closest"]["prim"][i]) if g["closest"]["prim"][i] != 0xFFFFFFFF else j
        tri = np.ascontiguousarray(flat[j])
        want = orc.intersect_tri(tri, o[i], d[i])
        for fast in (0, 1):
            r = h.host_mt(tri.ctypes.data, np.ascontiguousarray(o[i]).ctypes.data,
                          np.ascontiguousarray(d[i]).ctypes.data, fast, tuv.ctypes.data)
            assert r == want[0]
            if r:
                assert tuv.view(np.uint32).tolist() == np.array(want[1:], np.float32).view(np.uint32).tolist()


def test_host_header_mt_fast_decides_like_oracle_on_adversarial_pairs():
    """mt_fast's division-free bounds take the oracle's decision where they are tight (tests/mt_cases.py)"""
    h = host()
    tuv = np.zeros(3, np.float32)
    hits = 0
    for tri, o, d in adversarial_mt_cases():
        want = orc.intersect_tri(tri, o, d)
        r = h.host_mt(tri.ctypes.data, o.ctypes.data, d.ctypes.data, 1, tuv.ctypes.data)
        assert r == want[0], (tri, o, d)
        if r:
            hits += 1
            assert tuv.view(np.uint32).tolist() == np.array(want[1:], np.float32).view(np.uint32).tolist()
    assert hits > 100
