"""Batched persistent launch over many scenes (hipGraph replay + launch-shape autotuner): every hit record
equals the per-scene traversal, which equals brute force over the reference test, for every launch shape."""
import numpy as np
import pytest

import vierkant_amd as vk
from parity import assert_hits_equal
from vierkant_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch(ctx):
    scenes = list(W.many_scenes(n_scenes=12, seed=3, tri_range=(200, 20000), ray_range=(500, 9000)))
    bvhs, rays, hits, flats = [], [], [], []
    for pos, idx, r in scenes:
        bvhs.append(ctx.bvh_build(pos, idx))
        rays.append(vk.DeviceArray.from_host(ctx, r))
        hits.append(vk.DeviceArray(ctx, len(r), vk.HIT_DTYPE))
        flats.append(W.flat_triangles(pos, idx))
    return scenes, bvhs, rays, hits, flats


def per_scene(ctx, small_batch):
    scenes, bvhs, rays, hits, flats = small_batch
    return [ctx.trace_closest(b, r) for (_, _, r), b in zip(scenes, bvhs)]


def test_batch_equals_per_scene_and_brute_force(ctx, small_batch):
    scenes, bvhs, rays, hits, flats = small_batch
    batch = vk.Batch(ctx, list(zip(bvhs, rays, hits)))
    assert batch.total_rays() == sum(len(r) for _, _, r in scenes)
    batch.launch()
    want = per_scene(ctx, small_batch)
    for s, (h, w) in enumerate(zip(hits, want)):
        assert_hits_equal(h.host(), w, f"scene {s}")
    # a subset of every scene against brute force
    for s, ((pos, idx, r), h) in enumerate(zip(scenes, hits)):
        sub = np.arange(0, len(r), 7)
        assert_hits_equal(h.host()[sub], ctx.brute_closest(flats[s], r[sub]), f"scene {s} vs brute force")


def test_batch_graph_replay_and_autotune(ctx, small_batch):
    scenes, bvhs, rays, hits, flats = small_batch
    batch = vk.Batch(ctx, list(zip(bvhs, rays, hits)))
    want = per_scene(ctx, small_batch)
    poison = np.full(max(h.nbytes for h in hits), 0xAB, np.uint8)
    for cfg in [(64, 0, 64, 1, 0), (128, 8, 256, 4, 4), (256, 32, 4096, 8, 0), (256, 16, 1024, 2, 4),
                (64, 64, 256, 16, 1), (256, 64, 1024, 2, 64), (128, 24, 512, 8, 2), (64, 48, 1024, 4, 8)]:
        batch.config = cfg
        assert batch.config == cfg
        for _ in range(2):                          # replays of the same captured graph
            for h in hits:                          # poison outputs between replays
                ctx.check(vk.lib().vkrt_copy(ctx.handle, h.ptr, poison.ctypes.data, h.nbytes, 0, 0))
            batch.launch()
            for s, (h, w) in enumerate(zip(hits, want)):
                assert_hits_equal(h.host(), w, f"cfg {cfg} scene {s}")
    best, ms = batch.autotune(repeats=1)
    assert ms > 0 and batch.config == best and best in vk.Batch.candidates()
    batch.launch()
    for s, (h, w) in enumerate(zip(hits, want)):
        assert_hits_equal(h.host(), w, f"autotuned {best} scene {s}")
    for bad in [(96, 16, 1024, 8, 0), (256, 12, 1024, 8, 0), (256, 16, 1000, 8, 0), (256, 16, 1024, 0, 0),
                (256, 16, 1024, 8, 65)]:
        with pytest.raises(vk.VkrtError):
            batch.config = bad


def test_batch_closest_one_shot_and_empty_items(ctx, small_batch):
    scenes, bvhs, rays, hits, flats = small_batch
    empty_bvh = ctx.bvh_build(np.zeros((1, 3), np.float32), np.zeros(0, np.uint32))
    r0 = rays[0]
    h0 = vk.DeviceArray(ctx, len(r0), vk.HIT_DTYPE)
    z = vk.DeviceArray(ctx, 1, vk.HIT_DTYPE)
    items = (vk.BatchItem * 3)(vk.BatchItem(bvhs[0].handle, r0.ptr, hits[0].ptr, len(r0), 0),
                               vk.BatchItem(empty_bvh.handle, r0.ptr, h0.ptr, len(r0), 0),
                               vk.BatchItem(bvhs[1].handle, rays[1].ptr, z.ptr, 0, 0))
    ctx.check(vk.lib().vkrt_batch_closest(ctx.handle, items, 3, 0))
    assert_hits_equal(hits[0].host(), ctx.trace_closest(bvhs[0], scenes[0][2]), "one-shot")
    miss = h0.host()
    assert np.all(miss["prim"] == 0xFFFFFFFF) and np.all(np.isinf(miss["t"]))
