"""Comparison helpers for parity tests (bit-exact where the reference arithmetic is reproduced)."""
import numpy as np

INVALID = 0xFFFFFFFF


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_hits_equal(got, want, label=""):
    """closest-hit records identical: same prim, and t/u/v bit patterns equal (misses: t=inf, u=v=0)"""
    assert got.shape == want.shape
    bad = np.nonzero(got["prim"] != want["prim"])[0]
    assert len(bad) == 0, f"{label}: {len(bad)} prim mismatches, first {bad[:8]}: got {got[bad[:4]]} want {want[bad[:4]]}"
    for f in ("t", "u", "v"):
        diff = np.nonzero(bits(got[f]) != bits(want[f]))[0]
        assert len(diff) == 0, f"{label}: {len(diff)} {f} bit mismatches, first {diff[:8]}: {got[diff[:4]]} vs {want[diff[:4]]}"


def assert_tri_results_equal(got, want):
    assert np.array_equal(got["type"], want["type"]), f"{np.sum(got['type'] != want['type'])} type mismatches"
    for f in ("distance", "u", "v"):
        assert np.array_equal(bits(got[f]), bits(want[f])), f"{f} mismatches"


def assert_results_equal(got, want):
    assert np.array_equal(got["type"], want["type"]), f"{np.sum(got['type'] != want['type'])} type mismatches"
    diff = np.nonzero(bits(got["distance"]) != bits(want["distance"]))[0]
    assert len(diff) == 0, f"{len(diff)} distance mismatches: {got[diff[:4]]} vs {want[diff[:4]]}"
