"""The batched launch over the full many-scenes workload (256 scenes, log-uniform 1k-500k triangles and
16k-2M rays each, 110,795,658 rays, seed 0) under every launch shape the autotuner can pick, checked bit for
bit (t, u, v, prim) against the oracle's brute force over the reference test (src/intersection.cpp:114-129)
on a fixed block of rays of every scene. Every ray of that sample also passes the traversal's culling
criterion (the winner's own padded box is entered at the winner's t), and every scene's BVH passes the host
invariant check, so the agreement is the one the culling argument predicts, not a chance one."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from parity import assert_hits_equal
from vierkant_amd import workloads as W

pytestmark = pytest.mark.gpu

BLOCK = 48  # contiguous rays per scene checked


@pytest.fixture(scope="module")
def many(ctx):
    bvhs, rays, hits, sample = [], [], [], []
    n_rays = 0
    for s, (pos, idx, r) in enumerate(W.many_scenes(n_scenes=256)):
        bvh = ctx.bvh_build(pos, idx)
        check = bvh.validate(pos)
        assert check["ok"], f"scene {s}: BVH invariants {check}"
        bvhs.append(bvh)
        rays.append(vk.DeviceArray.from_host(ctx, r))
        hits.append(vk.DeviceArray(ctx, len(r), vk.HIT_DTYPE))
        first = int(np.random.default_rng(1000 + s).integers(0, len(r) - BLOCK))
        want, safe = orc.brute_closest_cull_safe(W.flat_triangles(pos, idx), r[first:first + BLOCK])
        assert safe.all(), f"scene {s}: rays {np.nonzero(safe == 0)[0] + first} fail the culling criterion"
        sample.append((first, want))
        n_rays += len(r)
    assert n_rays == 110795658
    batch = vk.Batch(ctx, list(zip(bvhs, rays, hits)))
    yield batch, hits, sample
    batch.close()
    for a in rays + hits:
        a.free()
    for b in bvhs:
        b.close()


def _check(ctx, batch, hits, sample, label):
    poison = np.full(BLOCK * vk.HIT_DTYPE.itemsize, 0xAB, np.uint8)
    out = np.empty(BLOCK, vk.HIT_DTYPE)
    for h, (first, _) in zip(hits, sample):
        ctx.check(vk.lib().vkrt_copy(ctx.handle, h.ptr + first * vk.HIT_DTYPE.itemsize, poison.ctypes.data,
                                     poison.nbytes, 0, 0))
    batch.launch()
    for s, (h, (first, want)) in enumerate(zip(hits, sample)):
        ctx.check(vk.lib().vkrt_copy(ctx.handle, out.ctypes.data, h.ptr + first * vk.HIT_DTYPE.itemsize,
                                     out.nbytes, 1, 0))
        assert_hits_equal(out, want, f"{label} scene {s} rays {first}..{first + BLOCK}")


def test_every_candidate_shape_matches_the_oracle(ctx, many):
    batch, hits, sample = many
    shapes = vk.Batch.candidates()
    assert len(shapes) >= 100
    assert {c[1] for c in shapes} == {16, 32, 48, 64}
    for cfg in shapes:
        batch.config = cfg
        _check(ctx, batch, hits, sample, f"shape {cfg}")


def test_autotuned_shape_matches_the_oracle(ctx, many):
    batch, hits, sample = many
    batch.config = (256, 16, 1024, 8, 4)
    best, ms = batch.autotune(repeats=1)
    assert best in vk.Batch.candidates() and ms > 0
    print(f"autotuned shape {best}: {ms:.2f} ms per launch", flush=True)
    for replay in range(3):
        _check(ctx, batch, hits, sample, f"autotuned {best} replay {replay}")
