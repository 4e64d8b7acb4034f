"""Headline benchmark: closest-hit rays per second through the BVH on one MI355X.

Workload (closest-hit on the 1.8M-triangle scene): a 512x512 fBm heightfield plus 1,000 subdivision-3
icospheres (1,804,288 triangles), 4,194,304 pinhole rays at 2048x2048 and 4,194,304 cosine-hemisphere
bounce rays from their hits. One step traces both ray sets through the BVH (8,388,608 closest-hit rays);
rays, BVH and hit buffers are resident in HBM before the timed region. The 4,194,304 shadow rays are
traced with the any-hit kernel and reported beside the headline.

Prints one JSON line (rank 0). Options: --gpus (only 1 is supported), --steps, --warmup,
--cpu-seconds (budget of the CPU baseline sample), --no-cpu-baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_HBM_GBPS = 8000.0
CLOSEST_BYTES_PER_RAY = 48   # 32 B ray in + 16 B hit out
MAX_LEAF = 2                 # triangles per BVH leaf (the library default; fastest with the 4-wide layout)
ANY_BYTES_PER_RAY = 33       # 32 B ray in + 1 B flag out


# wave64 VALU instructions per second: 1,024 SIMDs of 32 FP32 lanes (157.3 TFLOP/s = 256 CUs x 4 SIMDs x 32
# lanes x 2 flop x 2.4 GHz), 2 cycles per instruction; the brute-force kernel sustains 8.1e11/s, above what
# 16-lane SIMDs (4 cycles, 6.1e11/s) could issue
MI355X_VALU_ISSUE = 1024 * 2.4e9 / 2


def profiled(kernel, field):
    """`field` of `kernel` from the committed rocprofv3 summary (profiles/k_trace_heightfield_spheres.json:
    HBM bytes per launch from FETCH_SIZE / WRITE_SIZE, each scaled by its plain-copy calibration; VALU
    instructions per launch from SQ_INSTS_VALU), or None"""
    path = os.path.join(ROOT, "profiles", "k_trace_heightfield_spheres.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data["kernels"][kernel][field], data["source"]
    except (OSError, KeyError, ValueError):
        return None, None


def profiled_traffic(kernel):
    return profiled(kernel, "bytes_per_launch")


def rays_per_step_of(primary, bounce):
    return len(primary) + len(bounce)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(tris, rays, budget_s):
    """the oracle's brute-force closest hit (restatement of src/intersection.cpp:114-129 in a loop), timed
    on this host on a bounded random subset of the primary + bounce rays, extrapolated to rays/s: on every
    thread this process may use (std::thread::hardware_concurrency of the reference's harness) and on one
    thread, one warm-up and the median of 5 runs each"""
    import numpy as np
    orc = oracle()
    nthreads, model = host_threads()
    rng = np.random.default_rng(0)
    probe = rays[rng.choice(len(rays), 256, replace=False)]
    t0 = time.perf_counter()
    orc.brute_closest(tris, probe, nthreads)
    dt = time.perf_counter() - t0
    # about budget_s / 6 seconds per multi-threaded run (1 warm-up + 5 timed)
    n = int(min(32768, max(256, 256 * budget_s / 6.0 / max(dt, 1e-6))))
    sample = rays[np.sort(rng.choice(len(rays), n, replace=False))]
    times = []
    for k in range(6):
        t0 = time.perf_counter()
        orc.brute_closest(tris, sample, nthreads)
        if k:
            times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    one = sample[: max(16, n // (4 * nthreads))]
    times1 = []
    for k in range(3):
        t0 = time.perf_counter()
        orc.brute_closest(tris, one, 1)
        if k:
            times1.append(time.perf_counter() - t0)
    dt1 = float(np.median(times1))
    return {"value": n / dt / 1e9, "unit": "Grays/s", "cores": nthreads, "kind": "port",
            "cpu_model": model, "one_thread_grays_per_s": len(one) / dt1 / 1e9,
            "sample": f"{n} random closest-hit rays of the workload, brute force over all {len(tris)} triangles "
                      f"({n * len(tris):.3e} ray-triangle tests, median {dt:.2f} s of 5 runs on {nthreads} threads; "
                      f"{len(one)} rays on 1 thread, median {dt1:.2f} s of 2), extrapolated to rays/s"}


def brute_force_leg(ctx, reps=10):
    """the brute-force closest hit of 65,536 rays (origins U[-3,3]^3, directions uniform on S^2, seed 0)
    against the 20,480-triangle jittered icosphere (subdivision 5, radius 1, vertices N(0, 0.01)): one
    ray-triangle test per pair, 1,342,177,280 per run, 52 FP32 operations each on the full path of
    src/intersection.cpp:114-129 (no FMA). Timed on the GPU (HBM-resident arrays, hipEvents around `reps`
    runs) and, as the CPU baseline of this workload, the oracle's full run on every host thread (1 warm-up,
    median of 5) and a 4,096-ray sample on one thread. The GPU records are checked bit for bit against the
    oracle's full run."""
    import numpy as np
    import vierkant_amd as vk
    from vierkant_amd import workloads as W
    orc = oracle()
    pos, idx, rays = W.icosphere_rays()
    tris = W.flat_triangles(pos, idx)
    n_tests = len(rays) * len(tris)
    d_tris, d_rays = vk.DeviceArray.from_host(ctx, tris), vk.DeviceArray.from_host(ctx, rays)
    d_hits = vk.DeviceArray(ctx, len(rays), vk.HIT_DTYPE)
    # untimed calls first (about 20 ms of work): the clock is still rising during the first few ms of a
    # busy GPU, and the timed calls should see the steady state
    for _ in range(reps + 4):
        ctx.brute_closest(d_tris, d_rays, out=d_hits, flags=vk.ASYNC)
    ctx.sync()
    ev = vk.HipEvents(2)
    ev.record(0, ctx.stream)
    for _ in range(reps):
        ctx.brute_closest(d_tris, d_rays, out=d_hits, flags=vk.ASYNC)
    ev.record(1, ctx.stream)
    ev.sync(1)
    ms = ev.ms(0, 1) / reps
    got = d_hits.host()
    for a in (d_tris, d_rays, d_hits):
        a.free()
    nthreads, model = host_threads()
    cpu = []
    for k in range(6):
        t0 = time.perf_counter()
        want = orc.brute_closest(tris, rays, nthreads)
        if k:
            cpu.append(time.perf_counter() - t0)
    cpu_s = float(np.median(cpu))
    one = rays[:4096]
    t0 = time.perf_counter()
    orc.brute_closest(tris, one, 1)
    one_s = time.perf_counter() - t0
    mismatch = int(np.sum((got["prim"] != want["prim"]) | (got["t"].view(np.uint32) != want["t"].view(np.uint32)) |
                          (got["u"].view(np.uint32) != want["u"].view(np.uint32)) |
                          (got["v"].view(np.uint32) != want["v"].view(np.uint32))))
    flops = 52.0 * n_tests
    tflops = flops / (ms * 1e-3) / 1e12
    return {"workload": "brute-force closest hit, 65,536 rays x 20,480-triangle jittered icosphere (subdiv 5, "
                        "r 1, N(0, 0.01)), seed 0", "tests_per_run": n_tests,
            "gpu_ms": round(ms, 4), "gpu_tests_per_s": round(n_tests / (ms * 1e-3), 1),
            "gpu_rays_per_s": round(len(rays) / (ms * 1e-3), 1),
            "roofline": {"bound": "fp32 valu", "kernel": "k_brute_closest (whole call: prep + fill + test + finalize)",
                         "achieved": round(tflops, 3), "peak": 157.3, "unit": "TFLOP/s",
                         "frac": round(tflops / 157.3, 5), "flop_per_test": 52, "flop_per_run": flops},
            "verified_vs_oracle_full_run": {"rays": len(rays), "mismatches": mismatch,
                                            "hits": int(np.sum(want["prim"] != 0xFFFFFFFF))},
            "cpu_full_run": {"threads": nthreads, "cpu_model": model, "median_s": round(cpu_s, 3),
                             "tests_per_s": round(n_tests / cpu_s, 1), "rays_per_s": round(len(rays) / cpu_s, 1),
                             "one_thread_tests_per_s": round(len(one) * len(tris) / one_s, 1),
                             "kind": "port"},
            "gpu_over_cpu": round(cpu_s / (ms * 1e-3), 1)}


def path_tracer_frame(ctx, spp, lit=False, instanced=False):
    """one PBRPathTracer batch of the Cornell-style room (12 quads + 200 spheres + the jittered icosphere,
    276,504 triangles) at 1920x1080, max_trace_depth 8, `spp` samples per pixel; `lit` adds the room's six
    analytic lights (next-event estimation with a transmittance shadow ray per surface vertex, MIS);
    `instanced` traces the room as a two-level scene (202 placements of 10 meshes: the room's quads, 8 shared
    unit spheres under random rotations and scales, the jittered icosphere rotated and scaled), and also
    times the same placed triangles as a flat scene"""
    import vierkant_amd as vk
    from vierkant_amd import workloads as W
    s = vk.default_settings()
    s.resolution[0], s.resolution[1] = 1920, 1080
    s.num_samples, s.max_trace_depth = spp, 8
    pt = vk.PathTracer(ctx, s, seed=0)
    if instanced:
        scene, mesh_first, inst = W.room_instanced()
        pt.set_instanced_scene(scene, mesh_first, inst)
    else:
        scene = W.room_scene()
        pt.set_scene(scene)
    if lit:
        pt.set_lights(W.room_lights())
    cam = W.room_camera()
    s_warm = vk.default_settings()
    s_warm.resolution[0], s_warm.resolution[1] = 1920, 1080
    s_warm.num_samples, s_warm.max_trace_depth = 1, 8
    pt.settings = s_warm
    pt.render(cam)                      # warm-up batch (1 spp)
    pt.reset_accumulator()
    pt.settings = s
    pt.render(cam)
    st = pt.stats()
    px = pt.pixels()
    seg_bytes = 332      # per path segment: SoA ray, hit and path state in/out, 3 indices + 3 vertices, material
    gbps = seg_bytes * st.segments / (st.raytrace_ms * 1e-3) / 1e9
    what = ("%d placements of %d meshes (%d mesh triangles, 276,504 placed)"
            % (len(inst), len(mesh_first) - 1, len(scene["indices"]) // 3)) if instanced else "276,504 triangles"
    out = {"workload": "1920x1080, max_trace_depth 8, %d spp, %s%s"
                       % (spp, what, ", 6 analytic lights (NEE + MIS)" if lit else ""),
           "frame_ms": round(st.raytrace_ms, 3), "samples": int(st.samples), "segments": int(st.segments),
           "samples_per_s": round(st.samples / (st.raytrace_ms * 1e-3), 1),
           "segments_per_s": round(st.segments / (st.raytrace_ms * 1e-3), 1),
           "segment_gbps": round(gbps, 2),
           "roofline": {"bound": "hbm", "achieved": round(gbps, 2), "peak": MI355X_HBM_GBPS, "unit": "GB/s",
                        "frac": round(gbps / MI355X_HBM_GBPS, 5), "bytes_per_segment": seg_bytes},
           "mean_radiance": float(px["radiance"].mean())}
    pt.close()
    if instanced:
        # the same placed triangles as one flat scene, same settings and seed
        flat = vk.PathTracer(ctx, s_warm, seed=0)
        flat.set_scene(W.flatten_instances(scene, mesh_first, inst))
        if lit:
            flat.set_lights(W.room_lights())
        flat.render(cam)
        flat.reset_accumulator()
        flat.settings = s
        flat.render(cam)
        fst = flat.stats()
        out["flat_equivalent"] = {"frame_ms": round(fst.raytrace_ms, 3), "segments": int(fst.segments),
                                  "instanced_over_flat": round(st.raytrace_ms / fst.raytrace_ms, 3)}
        flat.close()
        out["vs_cpu_integrator"] = pt_vs_integrator(ctx, scene, cam, w=240, h=135, spp=8,
                                                    instanced=(mesh_first, inst))
    elif not lit:
        out["vs_cpu_integrator"] = pt_vs_integrator(ctx, scene, cam)
    return out


def pt_vs_integrator(ctx, scene, cam, w=480, h=270, spp=16, instanced=None):
    """the GPU tracer and the CPU reference integrator (oracle, same RNG streams and arithmetic) on the same
    scene, camera and seed at equal samples per pixel, at a reduced size: RMSE relative to the mean and the
    share of pixels within 1e-4 (the integrator's one-ulp sensitivity bounds what two math libraries give)"""
    import numpy as np
    import vierkant_amd as vk
    from vierkant_amd import workloads as W
    orc = oracle()
    s = vk.default_settings()
    s.resolution[0], s.resolution[1] = w, h
    s.num_samples, s.max_trace_depth = spp, 8
    pt = vk.PathTracer(ctx, s, seed=0)
    if instanced is not None:
        pt.set_instanced_scene(scene, *instanced)
    else:
        pt.set_scene(scene)
    pt.render(cam)
    g = pt.pixels()["radiance"].astype(np.float64)
    pt.close()
    t0 = time.perf_counter()
    c = orc.pt_render(scene, cam, w, h, spp, 8, W.minstd_rand0(0, 1)[0],
                      instanced=instanced)[..., :3].astype(np.float64)
    cpu_s = time.perf_counter() - t0
    close = float(np.mean(np.abs(g - c).max(axis=2) <= 1e-4 * np.maximum(1.0, c.max(axis=2))))
    return {"size": [w, h], "spp": spp, "rmse": float(np.sqrt(np.mean((g - c) ** 2))),
            "rmse_rel_mean": float(np.sqrt(np.mean((g - c) ** 2)) / c.mean()), "share_within_1e-4": round(close, 5),
            "cpu_integrator_s": round(cpu_s, 2), "cpu_threads": host_threads()[0]}


def radiance_cache_run(ctx, ref_spp, max_frames, cache_depth=2, train_stride=512, max_train_records=4096):
    """time to equal RMSE on the room scene at 1920x1080, max_trace_depth 8. The target is plain path
    tracing at 64 spp, timed as one 64-spp batch (its fastest form). The reference image is plain path
    tracing at ref_spp. Cache frames trace 1 spp: paths that are not training paths stop at hit depth
    `cache_depth` (after two bounces) and add beta * cache; one pixel in `train_stride` traces a
    full-length training path, and the cache trains once per frame on at most `max_train_records` of their
    vertices. Frame times are GPU event times
    (tracing, cache queries and training); image read-back and RMSE are outside them."""
    import numpy as np
    import vierkant_amd as vk
    from vierkant_amd import workloads as W
    scene = W.room_scene()
    cam = W.room_camera()

    def settings(spp, batches):
        st = vk.default_settings()
        st.resolution[0], st.resolution[1] = 1920, 1080
        st.num_samples, st.max_trace_depth, st.max_num_batches = spp, 8, batches
        return st

    # the reference: ref_spp / 64 batches of 64 spp (max_num_batches 0: each render replaces the image with
    # a fresh batch under the next per-frame seed), accumulated in float64 on the host
    t0 = time.perf_counter()
    ref_pt = vk.PathTracer(ctx, settings(64, 0), seed=1000)
    ref_pt.set_scene(scene)
    n_ref = max(1, ref_spp // 64)
    ref = np.zeros((1080, 1920, 3), np.float64)
    for _ in range(n_ref):
        ref_pt.render(cam)
        ref += ref_pt.pixels()["radiance"]
    ref /= n_ref
    ref_pt.close()
    ref_s = time.perf_counter() - t0

    def rmse(img):
        return float(np.sqrt(np.mean((img.astype(np.float64) - ref) ** 2)))

    plain = vk.PathTracer(ctx, settings(64, 0), seed=0)
    plain.set_scene(scene)
    # max_num_batches 0 keeps the batch index at 0 (min(index + 1, 0), PBRPathTracer.cpp:211), so every
    # render replaces the image: the first warms the kernels, the second is the timed 64-spp image
    plain.render(cam)
    t0 = time.perf_counter()
    plain.render(cam)
    plain_wall_ms = (time.perf_counter() - t0) * 1e3
    assert plain.current_batch() == 0
    target = rmse(plain.pixels()["radiance"])
    plain_ms = float(plain.stats().raytrace_ms)
    plain.close()

    # code objects of the cache kernels load on their first launch: warm them on a throwaway tracer
    warm = vk.PathTracer(ctx, settings(1, 0), seed=0)
    warm.set_scene(scene)
    warm_cache = vk.RadianceCache(ctx, seed=7)
    warm.set_radiance_cache(warm_cache, cache_depth=cache_depth, train_stride=train_stride,
                            max_train_records=max_train_records)
    warm.render(cam)
    warm.close()
    warm_cache.close()

    pt = vk.PathTracer(ctx, settings(1, 1 << 20), seed=0)
    pt.set_scene(scene)
    cache = vk.RadianceCache(ctx, seed=1)
    pt.set_radiance_cache(cache, cache_depth=cache_depth, train_stride=train_stride,
                          max_train_records=max_train_records)
    t_ms, wall_ms, hit, curve, frame_ms = 0.0, 0.0, None, [], []
    for k in range(1, max_frames + 1):
        t0 = time.perf_counter()
        pt.render(cam)  # synchronous: the host's wall time of the whole frame (launches, queries, training)
        wall_ms += (time.perf_counter() - t0) * 1e3
        frame_ms.append(float(pt.stats().raytrace_ms))
        t_ms += frame_ms[-1]
        r = rmse(pt.pixels()["radiance"])
        if k in (1, 2, 4, 8, 16, 32, 64, 128) or (hit is None and r <= target):
            curve.append([k, round(t_ms, 2), round(wall_ms, 2), round(r, 5)])
        if hit is None and r <= target:
            hit = (k, t_ms, wall_ms)
            break
    pt.close()
    kernels = cache_kernel_rooflines(ctx, cache)
    cache.close()
    return {"workload": "room scene (276,504 triangles), 1920x1080, max_trace_depth 8; target = RMSE of plain "
                        f"path tracing at 64 spp against a {ref_spp}-spp plain reference",
            "target_rmse": round(target, 5), "plain_64spp_ms": round(plain_ms, 3),
            "cache_depth": cache_depth, "train_stride": train_stride, "max_train_records": max_train_records,
            "reference": f"{n_ref} batches of 64 spp accumulated in float64",
            "cache_frames_to_target": None if hit is None else hit[0],
            "cache_ms_to_target": None if hit is None else round(hit[1], 3),
            "speedup": None if hit is None else round(plain_ms / hit[1], 3),
            "plain_64spp_wall_ms": round(plain_wall_ms, 3),
            "cache_wall_ms_to_target": None if hit is None else round(hit[2], 3),
            "speedup_wall": None if hit is None else round(plain_wall_ms / hit[2], 3),
            "cache_frame_ms_mean": round(float(np.mean(frame_ms)), 3),
            "rmse_curve_frames_gpu_ms_wall_ms_rmse": curve, "reference_s": round(ref_s, 2),
            "kernels": kernels}


def cache_kernel_rooflines(ctx, cache, n_query=1 << 20, n_train=1 << 16, reps=10):
    """k_rc_query and one training step (forward, backward, gradient scatter and Adam) timed alone with
    hipEvents on HBM-resident records, against the dense BF16 matrix-core peak (2.5 PFLOP/s): 43,008 FLOP
    per query (64-wide input, five 64x64 layers, 64 -> 16 output) and 3 x that per training record (forward
    and two backward products), 512 B of table reads per query"""
    import numpy as np
    import vierkant_amd as vk
    rng = np.random.default_rng(9)
    def records(n):
        # pos.xyz, roughness | wo.xyz, metalness | normal.xyz, transmission | albedo.rgb, - (the record layout)
        r = np.zeros((n, 16), np.float32)
        r[:, 0:3] = rng.uniform(0.0, 1.0, (n, 3))
        r[:, 3] = rng.uniform(0.05, 1.0, n)
        for k in (4, 8):
            d = rng.normal(size=(n, 3))
            r[:, k:k + 3] = d / np.linalg.norm(d, axis=1, keepdims=True)
        r[:, 7] = rng.uniform(0.0, 1.0, n) < 0.3
        r[:, 12:15] = rng.uniform(0.0, 1.0, (n, 3))
        return r
    q = vk.DeviceArray.from_host(ctx, records(n_query))
    qo = vk.DeviceArray(ctx, (n_query, 4), np.float32)
    t = vk.DeviceArray.from_host(ctx, records(n_train))
    tg = vk.DeviceArray.from_host(ctx, rng.uniform(0.0, 2.0, (n_train, 4)).astype(np.float32))
    ev = vk.HipEvents(2)
    lib, loss = vk.lib(), None
    import ctypes
    loss = ctypes.c_float()
    flags = vk.DEVICE_PTRS | vk.ASYNC
    ctx.check(lib.vkrt_rc_query(cache.handle, q.ptr, n_query, qo.ptr, flags))
    ev.record(0, ctx.stream)
    for _ in range(reps):
        ctx.check(lib.vkrt_rc_query(cache.handle, q.ptr, n_query, qo.ptr, flags))
    ev.record(1, ctx.stream)
    ev.sync(1)
    q_ms = ev.ms(0, 1) / reps
    ctx.check(lib.vkrt_rc_train(cache.handle, t.ptr, tg.ptr, n_train, flags, ctypes.byref(loss)))
    ctx.sync()
    ev.record(0, ctx.stream)
    for _ in range(reps):
        ctx.check(lib.vkrt_rc_train(cache.handle, t.ptr, tg.ptr, n_train, flags, ctypes.byref(loss)))
    ev.record(1, ctx.stream)
    ev.sync(1)
    t_ms = ev.ms(0, 1) / reps
    for a in (q, qo, t, tg):
        a.free()
    peak = 2500.0
    q_tf = 43008.0 * n_query / (q_ms * 1e-3) / 1e12
    t_tf = 3 * 43008.0 * n_train / (t_ms * 1e-3) / 1e12
    return {"query": {"records": n_query, "ms": round(q_ms, 4), "tflops": round(q_tf, 2),
                      "mfma_bf16_frac": round(q_tf / peak, 5),
                      "table_read_gbps": round(512.0 * n_query / (q_ms * 1e-3) / 1e9, 1)},
            "train_step": {"records": n_train, "ms": round(t_ms, 4), "tflops": round(t_tf, 2),
                           "mfma_bf16_frac": round(t_tf / peak, 5)},
            "peak_tflops": peak}


def instanced_frame(ctx, primary, bounce, verify_rays, reps=5, flat_mesh=None, builder=None):
    """the heightfield + spheres scene as a two-level scene (1 heightfield instance + 1,000 instances of
    one unit IcoSphere(1, 3)): the same primary and bounce rays through the instance-level walk; with
    flat_mesh (positions, indices) also the flat scene's BVH walked as one identity instance, which separates
    the cost of the walk's level switching from that of the two-level structure"""
    import numpy as np
    import vierkant_amd as vk
    from vierkant_amd import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    meshes, mats, mesh_of = W.heightfield_spheres_instanced()
    builder = vk.BUILDER_PLOC if builder is None else builder   # the headline's builder
    bvhs = [ctx.bvh_build(p, i, builder=builder, max_leaf_size=MAX_LEAF) for p, i in meshes]
    inst = np.array([vk.instance(m, k, custom_index=j) for j, (m, k) in enumerate(zip(mats, mesh_of))],
                    dtype=vk.INSTANCE_DTYPE)
    sc = vk.Scene(ctx, bvhs, inst)
    ev = vk.HipEvents(3)
    dp, db = vk.DeviceArray.from_host(ctx, primary), vk.DeviceArray.from_host(ctx, bounce)
    hp = vk.DeviceArray(ctx, len(primary), vk.INSTANCE_HIT_DTYPE)
    hb = vk.DeviceArray(ctx, len(bounce), vk.INSTANCE_HIT_DTYPE)
    sc.trace_closest(dp, out=hp)
    ts = []
    for _ in range(reps):
        ev.record(0, ctx.stream)
        sc.trace_closest(dp, out=hp, flags=vk.ASYNC)
        ev.record(1, ctx.stream)
        sc.trace_closest(db, out=hb, flags=vk.ASYNC)
        ev.record(2, ctx.stream)
        ev.sync(2)
        ts.append((ev.ms(0, 1), ev.ms(1, 2)))
    p_ms, b_ms = float(np.median([t[0] for t in ts])), float(np.median([t[1] for t in ts]))
    one = None
    if flat_mesh is not None:
        fb = ctx.bvh_build(*flat_mesh, builder=builder, max_leaf_size=MAX_LEAF)
        ident = np.array([vk.instance(np.hstack([np.eye(3), np.zeros((3, 1))]), 0)], dtype=vk.INSTANCE_DTYPE)
        sc1 = vk.Scene(ctx, [fb], ident)
        sc1.trace_closest(dp, out=hp)
        t1 = []
        for _ in range(reps):
            ev.record(0, ctx.stream)
            sc1.trace_closest(dp, out=hp, flags=vk.ASYNC)
            ev.record(1, ctx.stream)
            sc1.trace_closest(db, out=hb, flags=vk.ASYNC)
            ev.record(2, ctx.stream)
            ev.sync(2)
            t1.append((ev.ms(0, 1), ev.ms(1, 2)))
        one = {"primary_ms": round(float(np.median([t[0] for t in t1])), 4),
               "bounce_ms": round(float(np.median([t[1] for t in t1])), 4)}
        sc1.close()
        fb.close()
        sc.trace_closest(dp, out=hp)
        sc.trace_closest(db, out=hb)
    mismatches = None
    if verify_rays:
        import oracle_binding as orc
        sel = np.sort(np.random.default_rng(3).choice(len(primary), verify_rays, replace=False))
        tris = [W.flat_triangles(p, i) for p, i in meshes]
        mismatches = 0
        for rays, dh in ((primary, hp), (bounce, hb)):
            got = dh.host()[sel]
            want = orc.scene_closest(tris, inst, rays[sel])
            mismatches += int(np.sum((got["prim"] != want["prim"]) | (got["instance"] != want["instance"]) |
                                     (got["t"].view(np.uint32) != want["t"].view(np.uint32))))
    n_inst, n_nodes, stack = sc.info()
    for a in (dp, db, hp, hb):
        a.free()
    sc.close()
    for b in bvhs:
        b.close()
    return {"workload": f"heightfield + 1,000 spheres as {n_inst} instances of 2 meshes "
                        f"({sum(len(i) // 3 for _, i in meshes)} mesh triangles, "
                        f"{'PLOC' if builder == vk.BUILDER_PLOC else 'LBVH'} mesh BVHs); same primary and bounce rays",
            "grays_per_s": round((len(primary) + len(bounce)) / ((p_ms + b_ms) * 1e-3) / 1e9, 4),
            "primary_ms": round(p_ms, 4), "bounce_ms": round(b_ms, 4), "top_level_nodes": n_nodes,
            "stack_bound": stack, "flat_bvh_as_one_instance": one,
            "verified_vs_instance_brute_force": None if mismatches is None else
            {"rays": 2 * verify_rays, "mismatches": mismatches}}


def oracle():
    """the CPU oracle binding (test infrastructure: used here only to check results and for the CPU baseline)"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as orc
    return orc


def host_threads():
    """threads this process may run on (the CPU share of the box), and the machine's model name"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def record(h):
    return {"t": float(h["t"]), "u": float(h["u"]), "v": float(h["v"]), "prim": int(h["prim"]),
            "t_bits": int(h["t"].view("u4"))}


def compare_with_oracle(got, want, safe, rays, limit, **where):
    """mismatching rays (prim or t bits differ) as diagnostic dicts, at most `limit`; plus the count"""
    import numpy as np
    bad = np.nonzero((got["prim"] != want["prim"]) | (got["t"].view(np.uint32) != want["t"].view(np.uint32)))[0]
    diags = []
    for i in bad[:limit]:
        r = rays[i]
        diags.append(dict(where, sample=int(i), ray=[float(x) for x in (r["ox"], r["oy"], r["oz"], r["tmin"],
                                                                        r["dx"], r["dy"], r["dz"], r["tmax"])],
                          got=record(got[i]), oracle=record(want[i]), winner_cull_safe=int(safe[i])))
    return len(bad), diags


def batched_frame(ctx, n_scenes, verify_per_scene):
    """256 independent scenes (log-uniform 1k-500k triangles, 16k-2M rays each) traced by one persistent
    launch replayed as a hipGraph, after the launch-shape autotuner picked block size, LDS stack depth,
    tile, blocks per CU and XCD chunk. `verify_per_scene` random rays of every scene are checked against the
    oracle's brute force (bit-exact t and prim); every mismatch is written out with the scene, the ray, both
    records, the launch shape and the host-side invariant check of that scene's BVH."""
    import numpy as np
    import vierkant_amd as vk
    from vierkant_amd import workloads as W
    orc = oracle()
    bvhs, rays, hits, host = [], [], [], []
    n_tri = 0
    t_build = 0.0
    bvh_bad = []
    for pos, idx, r in W.many_scenes(n_scenes=n_scenes):
        t0 = time.perf_counter()
        bvhs.append(ctx.bvh_build(pos, idx))
        t_build += time.perf_counter() - t0
        check = bvhs[-1].validate(pos)  # host-side invariants of every BVH (outside the timed launches)
        if not check["ok"]:
            bvh_bad.append(dict(check, scene=len(host)))
        rays.append(vk.DeviceArray.from_host(ctx, r))
        hits.append(vk.DeviceArray(ctx, len(r), vk.HIT_DTYPE))
        n_tri += len(idx) // 3
        sub = np.sort(np.random.default_rng(len(host)).choice(len(r), min(verify_per_scene, len(r)), replace=False))
        host.append((pos, W.flat_triangles(pos, idx), r[sub], sub))
    batch = vk.Batch(ctx, list(zip(bvhs, rays, hits)))
    total = batch.total_rays()
    default_cfg = batch.config
    batch.launch()
    ev = vk.HipEvents(2)

    def timed(k):
        ev.record(0, ctx.stream)
        for _ in range(k):
            batch.launch(flags=vk.ASYNC)
        ev.record(1, ctx.stream)
        ev.sync(1)
        return ev.ms(0, 1) / k
    default_ms = timed(3)
    t_tune = time.perf_counter()
    best, tune_ms = batch.autotune(repeats=2)
    tune_s = time.perf_counter() - t_tune
    tuned_ms = timed(5)
    # verification against the CPU oracle (outside the timed launches)
    n_checked = n_bad = n_safe = 0
    diags = []
    t_orc = time.perf_counter()
    for s, ((pos, flat, sub_rays, sub), h) in enumerate(zip(host, hits)):
        got = h.host()[sub]
        want, safe = orc.brute_closest_cull_safe(flat, sub_rays)
        n_checked += len(sub)
        n_safe += int(safe.sum())
        nb, d = compare_with_oracle(got, want, safe, sub_rays, 8 - len(diags), scene=s, shape=list(best))
        if nb:
            check = bvhs[s].validate(pos)
            for x in d:
                x["ray_index"] = int(sub[x["sample"]])
                x["bvh_check"] = check
            diags += d
        n_bad += nb
    orc_s = time.perf_counter() - t_orc
    for a in rays + hits:
        a.free()
    batch.close()
    for b in bvhs:
        b.close()
    achieved = CLOSEST_BYTES_PER_RAY * total / (tuned_ms * 1e-3) / 1e9
    return {"workload": f"{n_scenes} scenes, {n_tri} triangles, {total} closest-hit rays in one graph launch",
            "grays_per_s": round(total / (tuned_ms * 1e-3) / 1e9, 4), "launch_ms": round(tuned_ms, 3),
            "default_config": list(default_cfg), "default_launch_ms": round(default_ms, 3),
            "autotuned_config": {"block_size": best[0], "lds_stack_depth": best[1], "tile": best[2],
                                 "blocks_per_cu": best[3], "xcd_chunk": best[4]},
            "autotune_s": round(tune_s, 2), "candidate_shapes": len(vk.Batch.candidates()),
            "bvh_build_s": round(t_build, 3),
            "bvh_invariants": {"checked": n_scenes, "failed": len(bvh_bad), "first_failures": bvh_bad[:4]},
            "roofline": {"bound": "hbm", "kernel": "k_batch_trace", "achieved": round(achieved, 2),
                         "peak": MI355X_HBM_GBPS, "unit": "GB/s", "frac": round(achieved / MI355X_HBM_GBPS, 5),
                         "algorithmic_bytes_per_launch": CLOSEST_BYTES_PER_RAY * int(total)},
            "verified_vs_oracle": {"rays": n_checked, "mismatches": n_bad, "winner_cull_safe": n_safe,
                                   "oracle_s": round(orc_s, 1), "first_mismatches": diags}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verify-rays", type=int, default=65536,
                    help="primary and bounce rays each checked against GPU brute force")
    ap.add_argument("--oracle-rays", type=int, default=256,
                    help="of those, primary and bounce rays each checked against the CPU oracle")
    ap.add_argument("--sweep", action="store_true", help="time every trace launch shape and exit")
    ap.add_argument("--traversal", choices=("default", "step", "postponed"), default="default",
                    help="loop of the 4-wide walk (default: the library's)")
    ap.add_argument("--xcd-chunk", type=int, default=-1,
                    help="consecutive ray blocks per XCD turn of the trace launch (-1: the library's, 0: linear)")
    ap.add_argument("--pt-spp", type=int, default=64, help="samples per pixel of the path-tracer frame (0: skip)")
    ap.add_argument("--builder", choices=("lbvh", "ploc"), default="ploc", help="BVH builder of the headline scene")
    ap.add_argument("--ploc-radius", type=int, default=0, help="PLOC search radius (0: the library default, 16)")
    ap.add_argument("--pt-lit-spp", type=int, default=16,
                    help="samples per pixel of the lit path-tracer frame (0: skip)")
    ap.add_argument("--pt-inst-spp", type=int, default=16,
                    help="samples per pixel of the path-tracer frame over the instanced room (0: skip)")
    ap.add_argument("--batch-scenes", type=int, default=256, help="scenes of the batched launch (0: skip)")
    ap.add_argument("--brute-force", type=int, default=1,
                    help="time the brute-force leg (65,536 rays x 20,480 triangles) and its full CPU run (0: skip)")
    ap.add_argument("--batch-verify", type=int, default=256,
                    help="random rays per scene of the batched launch checked against the CPU oracle")
    ap.add_argument("--instanced", type=int, default=1, help="trace the two-level version of the scene (0: skip)")
    ap.add_argument("--rc-ref-spp", type=int, default=4096,
                    help="samples per pixel of the radiance-cache leg's reference image (0: skip the leg)")
    ap.add_argument("--rc-frames", type=int, default=128, help="most cache frames the radiance-cache leg traces")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (rays shard with no exchange; "
              "see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import vierkant_amd as vk
    from vierkant_amd import workloads as W

    log("generating scene")
    pos, idx = W.heightfield_spheres_scene()
    tris = W.flat_triangles(pos, idx)
    primary = W.pinhole_rays(2048, 2048)
    ctx = vk.Context(0)
    stream = ctx.stream

    log("building BVH")
    t0 = time.perf_counter()
    builder = {"lbvh": vk.BUILDER_LBVH, "ploc": vk.BUILDER_PLOC}[args.builder]
    bvh = ctx.bvh_build(pos, idx, builder=builder, max_leaf_size=MAX_LEAF, ploc_radius=args.ploc_radius)
    build_wall_ms = (time.perf_counter() - t0) * 1e3
    info = bvh.info()

    d_primary = vk.DeviceArray.from_host(ctx, primary)
    d_hits_p = vk.DeviceArray(ctx, len(primary), vk.HIT_DTYPE)
    ctx.trace_closest(bvh, d_primary, out=d_hits_p)
    hits_p = d_hits_p.host()
    log("generating secondary rays")
    bounce, shadow = W.secondary_rays(primary, hits_p, tris)
    d_bounce = vk.DeviceArray.from_host(ctx, bounce)
    d_shadow = vk.DeviceArray.from_host(ctx, shadow)
    d_hits_b = vk.DeviceArray(ctx, len(bounce), vk.HIT_DTYPE)
    d_occ = vk.DeviceArray(ctx, len(shadow), np.uint8)

    # parity check of a subset against GPU brute force (not timed)
    verified = None
    if args.verify_rays:
        log("verifying a subset against brute force")
        ctx.trace_closest(bvh, d_bounce, out=d_hits_b)
        hits_b = d_hits_b.host()
        sel = np.sort(np.random.default_rng(1).choice(len(primary), args.verify_rays, replace=False))
        mismatch = 0
        for rays_np, hits_np in ((primary, hits_p), (bounce, hits_b)):
            ref = ctx.brute_closest(tris, rays_np[sel])
            got = hits_np[sel]
            mismatch += int(np.sum((ref["prim"] != got["prim"]) |
                                   (ref["t"].view(np.uint32) != got["t"].view(np.uint32))))
        verified = {"vs_gpu_brute_force": {"rays": 2 * args.verify_rays, "mismatches": mismatch}}
        # a smaller sample against the CPU oracle itself, with the culling criterion of every winner
        orc = oracle()
        osel = sel[:: max(1, len(sel) // args.oracle_rays)][: args.oracle_rays]
        n_bad = n_safe = 0
        diags = []
        for name, rays_np, hits_np in (("primary", primary, hits_p), ("bounce", bounce, hits_b)):
            want, safe = orc.brute_closest_cull_safe(tris, rays_np[osel])
            nb, d = compare_with_oracle(hits_np[osel], want, safe, rays_np[osel], 8, ray_set=name)
            n_bad += nb
            n_safe += int(safe.sum())
            diags += d
        verified["vs_oracle"] = {"rays": 2 * len(osel), "mismatches": n_bad, "winner_cull_safe": n_safe,
                                 "first_mismatches": diags}

    if args.copy_mb:
        src = vk.DeviceArray(ctx, args.copy_mb << 20, np.uint8)
        dst = vk.DeviceArray(ctx, args.copy_mb << 20, np.uint8)
        ctx.check(vk.lib().vkrt_copy(ctx.handle, dst.ptr, src.ptr, args.copy_mb << 20, 2, 0))
        src.free()
        dst.free()

    flags = vk.ASYNC
    ev = vk.HipEvents(5)

    if args.traversal != "default":
        block, lds, _, _ = ctx.get_trace_config()
        ctx.set_trace_config(block, lds, vk.TRAVERSAL_POSTPONED if args.traversal == "postponed" else vk.TRAVERSAL_STEP)
    if args.xcd_chunk >= 0:
        block, lds, _, _ = ctx.get_trace_config()
        ctx.set_trace_config(block, lds, xcd_chunk=args.xcd_chunk)

    if args.sweep:
        for block, lds, trav, order in [(b, l, t, o) for o in (0, 1, 2, 4, 8, 16, 32, 64, 256)
                                        for t in (vk.TRAVERSAL_STEP, vk.TRAVERSAL_POSTPONED)
                                        for b in (64, 256) for l in (16,)]:
            ctx.set_trace_config(block, lds, trav, order)
            ctx.trace_closest(bvh, d_primary, out=d_hits_p)
            times = []
            for _ in range(5):
                ev.record(0, stream)
                ctx.trace_closest(bvh, d_primary, out=d_hits_p, flags=flags)
                ev.record(1, stream)
                ctx.trace_closest(bvh, d_bounce, out=d_hits_b, flags=flags)
                ev.record(2, stream)
                ev.sync(2)
                times.append((ev.ms(0, 1), ev.ms(1, 2)))
            p_ms, b_ms = np.median([t[0] for t in times]), np.median([t[1] for t in times])
            print(json.dumps({"block": block, "lds": lds, "traversal": trav, "xcd_chunk": order, "primary_ms": round(p_ms, 4),
                              "bounce_ms": round(b_ms, 4),
                              "grays_per_s": round(rays_per_step_of(primary, bounce) / ((p_ms + b_ms) * 1e-3) / 1e9, 4)}),
                  flush=True)
        return

    def step(timed):
        if timed:
            ev.record(0, stream)
        ctx.trace_closest(bvh, d_primary, out=d_hits_p, flags=flags)
        if timed:
            ev.record(1, stream)
        ctx.trace_closest(bvh, d_bounce, out=d_hits_b, flags=flags)
        if timed:
            ev.record(2, stream)

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step(False)
    ctx.sync()
    log(f"timing {args.steps} steps")
    prim_ms, bounce_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
        ev.sync(2)
        prim_ms.append(ev.ms(0, 1))
        bounce_ms.append(ev.ms(1, 2))
    ctx.sync()
    wall = time.perf_counter() - t0
    # steps are event-timed back to back on the stream; the host only reads the events between steps
    ms_per_step = float(np.mean(np.array(prim_ms) + np.array(bounce_ms)))
    rays_per_step = len(primary) + len(bounce)
    value = rays_per_step / (ms_per_step * 1e-3) / 1e9

    # any-hit shadow rays (reported beside the headline)
    any_ms = []
    for _ in range(max(3, args.steps // 2)):
        ev.record(3, stream)
        ctx.trace_any(bvh, d_shadow, out=d_occ, flags=flags)
        ev.record(4, stream)
        ev.sync(4)
        any_ms.append(ev.ms(3, 4))
    any_ms = float(np.mean(any_ms))

    # refit (the update path of a deforming mesh) with the same device-resident positions, after the timed
    # steps: boxes, node layout and the 4-wide boxes recomputed, the topology kept
    d_pos = vk.DeviceArray.from_host(ctx, pos)
    bvh.refit(d_pos)
    t0 = time.perf_counter()
    for _ in range(5):
        bvh.refit(d_pos)
    refit_ms = (time.perf_counter() - t0) * 1e3 / 5
    d_pos.free()

    launch_ms = float(np.mean(prim_ms + bounce_ms))
    traffic, traffic_source = profiled_traffic("k_trace_closest")
    valu, _ = profiled("k_trace_closest", "valu_insts_per_launch")
    achieved = CLOSEST_BYTES_PER_RAY * len(primary) / (launch_ms * 1e-3) / 1e9
    result = {
        "metric": "Grays/s closest-hit (1 GPU); time-to-equal-RMSE: radiance-cache vs plain PT",
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "closest-hit on the 1.8M-triangle scene: 512x512 fBm heightfield + 1,000 subdiv-3 "
                        "icospheres; 4,194,304 pinhole (2048x2048) + 4,194,304 cosine bounce rays per step",
            "n_tri": int(len(tris)),
            "rays_per_step": int(rays_per_step),
            "bvh": {"builder": args.builder, "ploc_radius": args.ploc_radius or 16, "max_leaf": MAX_LEAF, "nodes": int(info.n_nodes), "depth": int(info.depth),
                    "wide_nodes": int(info.n_wide_nodes), "wide_stack_bound": int(info.wide_stack),
                    "sah_cost": round(float(info.sah_cost), 2), "build_ms": round(float(info.build_ms), 2),
                    "build_wall_ms": round(build_wall_ms, 2), "refit_ms": round(refit_ms, 3)},
            "primary_ms": round(float(np.mean(prim_ms)), 4),
            "bounce_ms": round(float(np.mean(bounce_ms)), 4),
            "primary_grays_per_s": round(len(primary) / (np.mean(prim_ms) * 1e-3) / 1e9, 4),
            "bounce_grays_per_s": round(len(bounce) / (np.mean(bounce_ms) * 1e-3) / 1e9, 4),
            "any_hit_shadow_ms": round(any_ms, 4),
            "any_hit_grays_per_s": round(len(shadow) / (any_ms * 1e-3) / 1e9, 4),
            "verified_vs_brute_force": verified,
            "timed_wall_s": round(wall, 3),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_trace_closest",
            "achieved": round(achieved, 3),
            "peak": MI355X_HBM_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / MI355X_HBM_GBPS, 6),
            "traffic": None if traffic is None else round(traffic),
            "traffic_source": traffic_source,
            "algorithmic_bytes_per_launch": CLOSEST_BYTES_PER_RAY * len(primary),
            "avg_launch_ms": round(launch_ms, 4),
            # the second bound: VALU issue (the profile's instruction count per launch over this run's
            # mean launch time, against 1,024 SIMDs issuing one wave64 instruction per 2 cycles)
            "valu_issue": None if valu is None else {
                "insts_per_launch": round(valu), "issue_peak_per_s": MI355X_VALU_ISSUE,
                "frac": round(valu / (launch_ms * 1e-3) / MI355X_VALU_ISSUE, 4)},
        },
        "cpu_baseline": None,
    }
    ctx.sync()
    bvh.close()
    for a in (d_primary, d_hits_p, d_bounce, d_shadow, d_hits_b, d_occ):
        a.free()
    if args.instanced:
        log("two-level scene")
        result["config"]["instanced"] = instanced_frame(ctx, primary, bounce, 128 if args.verify_rays else 0,
                                                        flat_mesh=(pos, idx), builder=builder)
    if args.pt_spp:
        log("path tracer frame")
        result["config"]["path_tracer"] = path_tracer_frame(ctx, args.pt_spp)
    if args.pt_lit_spp:
        log("lit path tracer frame")
        result["config"]["path_tracer_lit"] = path_tracer_frame(ctx, args.pt_lit_spp, lit=True)
    if args.pt_inst_spp:
        log("instanced path tracer frame")
        result["config"]["path_tracer_instanced"] = path_tracer_frame(ctx, args.pt_inst_spp, instanced=True)
    if args.brute_force:
        log("brute force 65,536 x 20,480")
        result["config"]["brute_force"] = brute_force_leg(ctx)
    if args.batch_scenes:
        log("batched scenes")
        result["config"]["batched"] = batched_frame(ctx, args.batch_scenes, args.batch_verify)
    if args.rc_ref_spp:
        log("radiance cache: time to equal RMSE")
        result["config"]["radiance_cache"] = radiance_cache_run(ctx, args.rc_ref_spp, args.rc_frames)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        allrays = np.concatenate([primary, bounce])
        result["cpu_baseline"] = cpu_baseline(tris, allrays, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
