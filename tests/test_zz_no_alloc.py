"""The allocation-free closest-hit entry point (vkrt_intersect_brute_closest_noalloc), called through the
C ABI with device buffers, returns the same records as vkrt_intersect_brute_closest on two calls in a
row, the second with a smaller input, and rejects what it cannot take before launching anything."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from vierkant_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_TRI, MAX_RAY = 262144, 786432


def _case(subdiv, n_rays, seed):
    pos, idx = W.jittered_icosphere(subdiv=subdiv, seed=seed)
    tris = np.ascontiguousarray(W.flat_triangles(pos, idx))
    rng = np.random.default_rng(seed)
    o = rng.uniform(-3, 3, (n_rays, 3)).astype(np.float32)
    d = W.glm_normalize(rng.normal(size=(n_rays, 3)).astype(np.float32))
    rays = W.make_rays(o, d)
    # a few rays with a finite tmax and a positive tmin, so the interval ends take part
    rays["tmax"][::7] = np.float32(2.5)
    rays["tmin"][::11] = np.float32(1.0)
    return tris, rays


def _both(ctx, tris, rays):
    lib = vk.lib()
    n_tri, n_ray = len(tris), len(rays)
    d_tris = vk.DeviceArray.from_host(ctx, tris)
    d_rays = vk.DeviceArray.from_host(ctx, rays)
    d_ref = vk.DeviceArray(ctx, n_ray, vk.HIT_DTYPE)
    d_new = vk.DeviceArray(ctx, n_ray, vk.HIT_DTYPE)
    ctx.check(lib.vkrt_intersect_brute_closest(ctx.handle, d_tris.ptr, n_tri, d_rays.ptr, n_ray, d_ref.ptr,
                                               vk.DEVICE_PTRS))
    ctx.check(lib.vkrt_intersect_brute_closest_noalloc(ctx.handle, d_tris.ptr, n_tri, d_rays.ptr, n_ray,
                                                       d_new.ptr, vk.DEVICE_PTRS))
    ref, new = d_ref.host(), d_new.host()
    for a in (d_tris, d_rays, d_ref, d_new):
        a.free()
    return ref, new


def test_noalloc_equals_the_allocating_entry_point_twice(ctx):
    # the second call is smaller: whatever the first left in the shared array past its end is never read
    for subdiv, n_rays, seed in ((4, 40000, 0), (2, 3000, 1)):
        tris, rays = _case(subdiv, n_rays, seed)
        print(f"noalloc case: subdiv {subdiv}, {len(tris)} triangles, {n_rays} rays", flush=True)
        ref, new = _both(ctx, tris, rays)
        assert new.tobytes() == ref.tobytes()
        # and both are the oracle's records (the cases have hits: about 5 % of the rays)
        assert new.tobytes() == orc.brute_closest(tris, rays).tobytes()
        assert np.sum(ref["prim"] != vk.INVALID_PRIM) >= n_rays // 40


def test_noalloc_rejects_what_it_cannot_take(ctx):
    lib = vk.lib()
    dummy = vk.DeviceArray(ctx, 64, np.uint8)
    # oversize triangle and ray counts are refused before any pointer is read
    rc = lib.vkrt_intersect_brute_closest_noalloc(ctx.handle, dummy.ptr, MAX_TRI + 1, dummy.ptr, 1, dummy.ptr,
                                                  vk.DEVICE_PTRS)
    assert rc == vk.E_INVALID_ARG and b"VKRT_NOALLOC_MAX_TRI" in lib.vkrt_last_error(ctx.handle)
    rc = lib.vkrt_intersect_brute_closest_noalloc(ctx.handle, dummy.ptr, 1, dummy.ptr, MAX_RAY + 1, dummy.ptr,
                                                  vk.DEVICE_PTRS)
    assert rc == vk.E_INVALID_ARG and b"VKRT_NOALLOC_MAX_RAY" in lib.vkrt_last_error(ctx.handle)
    # host pointers would need staging buffers, i.e. an allocation
    host = np.zeros(64, np.uint8)
    rc = lib.vkrt_intersect_brute_closest_noalloc(ctx.handle, host.ctypes.data, 1, host.ctypes.data, 1,
                                                  host.ctypes.data, 0)
    assert rc == vk.E_INVALID_ARG and b"VKRT_DEVICE_PTRS" in lib.vkrt_last_error(ctx.handle)
    dummy.free()
    # the call after a rejection works
    tris, rays = _case(1, 500, 2)
    ref, new = _both(ctx, tris, rays)
    assert new.tobytes() == ref.tobytes()
