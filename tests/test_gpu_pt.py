"""GPU path tracer: parity with the CPU reference integrator (same RNG streams, same intersect and BSDF
arithmetic) at equal samples per pixel, closed-form scenes, and PBRPathTracer's batch / reset semantics
(src/PBRPathTracer.cpp:184-255, :677-682; shaders/slang/raypipeline.slang:249-270)."""
import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk
from vierkant_amd import workloads as W

pytestmark = pytest.mark.gpu


def settings(w, h, spp=1, depth=8, batches=0, suspend=1, suppress_reset=0, max_path_beta=0.0):
    s = vk.default_settings()
    s.resolution[0], s.resolution[1] = w, h
    s.num_samples, s.max_trace_depth, s.max_num_batches = spp, depth, batches
    s.suspend_trace_when_done, s.suppress_reset, s.max_path_beta = suspend, suppress_reset, max_path_beta
    return s


@pytest.fixture(scope="module")
def scene3():
    return W.room_scene()


def sphere_scene(material, radius=2.0):
    pos, idx = vk.icosphere(radius, 3)
    n = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    return {"positions": pos, "normals": n.astype(np.float32), "indices": idx,
            "tri_object": np.zeros(len(idx) // 3, np.uint32), "object_material": np.zeros(1, np.uint32),
            "materials": np.array([material], dtype=vk.MATERIAL_DTYPE), "environment": 1.0}


def empty_scene():
    return {"positions": np.zeros((0, 3), np.float32), "normals": np.zeros((0, 3), np.float32),
            "indices": np.zeros(0, np.uint32), "tri_object": np.zeros(0, np.uint32),
            "object_material": np.zeros(1, np.uint32), "materials": np.array([vk.material()], vk.MATERIAL_DTYPE),
            "environment": 1.0}


def test_pt_matches_cpu_integrator(ctx, scene3):
    """room scene, 160x90, 4 spp, depth 8: per-pixel radiance equals the CPU reference integrator up to
    transcendental-function ulps (ocml vs glibc cos/sin/pow), which can redirect a handful of paths"""
    w, h, spp = 160, 90, 4
    pt = vk.PathTracer(ctx, settings(w, h, spp), seed=0)
    pt.set_scene(scene3)
    cam = W.room_camera()
    pt.render(cam)
    got = pt.pixels()
    seed = W.minstd_rand0(0, 1)[0]
    want = orc.pt_render(scene3, cam, w, h, spp, 8, seed)
    g, c = got["radiance"], want[..., :3]
    assert np.isfinite(g).all()
    assert np.array_equal(got["alpha"], want[..., 3])
    assert np.array_equal(got["object_id"], want[..., 6].view(np.uint32))
    diff = np.abs(g - c).max(axis=2)
    close = np.mean(diff <= 1e-4 * np.maximum(1.0, c.max(axis=2)))
    rmse = float(np.sqrt(np.mean((g - c) ** 2)))
    print(f"GPU vs integrator: {close:.4f} of pixels within 1e-4, relative rmse {rmse / float(c.mean()):.5f}")
    assert close > 0.97, f"only {close:.4f} of pixels within 1e-4"
    assert rmse < 0.02 * float(c.mean()) + 1e-6, f"rmse {rmse} vs mean {c.mean()}"


def _alpha_room(scene3):
    """the room scene with every third material alpha-blended at 0.5 and every third (offset 1) alpha-masked
    at alpha 0.3, cutoff 0.5 (cut out): the trace and shadow kernels then take the any-hit filter"""
    sc = dict(scene3)
    mats = scene3["materials"].copy()
    mats["alpha_mode"][0::3] = vk.ALPHA_BLEND
    mats["color"][0::3, 3] = 0.5
    mats["alpha_mode"][1::3] = vk.ALPHA_MASK
    mats["color"][1::3, 3] = 0.3
    mats["alpha_cutoff"][1::3] = 0.5
    sc["materials"] = mats
    return sc


@pytest.mark.parametrize("alpha", [False, True])
@pytest.mark.parametrize("lit", [False, True])
def test_launch_shapes_give_the_same_image(ctx, scene3, lit, alpha):
    """per-depth launches as persistent grids over the live queue or as one lane per pixel at every depth, and
    the deep-bounce tail finished in one lane per path from depth 1, 2, 3 or 4 or never: the same image bit for
    bit and the same segment count (room scene, 960x540 so that the persistent grids are smaller than the full
    ones, 4 spp, depth 8; lit: the six analytic lights, next-event shadow rays included; alpha: alpha-blended
    and alpha-masked materials, so the any-hit filter runs in every kernel)"""
    sc = _alpha_room(scene3) if alpha else scene3
    shapes = [(vk.PT_LAUNCH_FULL_GRID, vk.PT_NO_TAIL), (vk.PT_LAUNCH_PERSISTENT, vk.PT_NO_TAIL),
              (vk.PT_LAUNCH_PERSISTENT, 1), (vk.PT_LAUNCH_FULL_GRID, 2), (vk.PT_LAUNCH_PERSISTENT, 3),
              (vk.PT_LAUNCH_FULL_GRID, vk.PT_TAIL_DEPTH_DEFAULT)]
    imgs = []
    for mode, tail in shapes:
        pt = vk.PathTracer(ctx, settings(960, 540, 4), seed=5)
        pt.set_scene(sc)
        pt.set_launch_mode(mode, tail)
        if lit:
            pt.set_lights(W.room_lights())
        pt.render(W.room_camera())
        imgs.append((pt.pixels(), pt.stats().segments))
        pt.close()
    ref = imgs[0]
    assert ref[1] > 960 * 540 * 4
    for (mode, tail), (px, seg) in zip(shapes[1:], imgs[1:]):
        assert px.tobytes() == ref[0].tobytes(), f"mode {mode}, tail {tail}"
        assert seg == ref[1], f"mode {mode}, tail {tail}: {seg} segments against {ref[1]}"
    bad = vk.PathTracer(ctx, settings(8, 8), seed=0)
    with pytest.raises(vk.VkrtError):
        bad.set_launch_mode(7)
    with pytest.raises(vk.VkrtError):
        bad.set_launch_mode(vk.PT_LAUNCH_PERSISTENT, 0)
    bad.close()


def test_pt_sky_only_bit_exact(ctx):
    """no geometry: every path misses at depth 0 and returns environment * environment_white(dir)"""
    w, h = 64, 48
    pt = vk.PathTracer(ctx, settings(w, h, 3), seed=7)
    pt.set_scene(empty_scene())
    cam = vk.make_camera(eye=(0.0, 1.0, 0.0), forward=(0.3, 0.4, -1.0), fov_y=1.2)
    pt.render(cam)
    got = pt.pixels()
    want = orc.pt_render(empty_scene(), cam, w, h, 3, 8, W.minstd_rand0(7, 1)[0])
    assert np.array_equal(got["radiance"], want[..., :3])
    assert np.all(got["alpha"] == 0) and np.all(got["object_id"] == 0xFFFF)


def test_pt_emission_closed_form(ctx):
    """camera inside a two-sided emissive sphere of black transmissive dielectric: the first hit adds
    emission * strength with beta = 1; every lobe of that BSDF is scaled by the black color, so beta
    becomes 0 and Russian roulette stops the path (q = 1)"""
    m = vk.material(color=(0.0, 0.0, 0.0), emission=(1.0, 0.5, 0.25), emission_strength=2.0, transmission=1.0,
                    two_sided=True)
    pt = vk.PathTracer(ctx, settings(40, 30, 2), seed=3)
    pt.set_scene(sphere_scene(m))
    pt.render(vk.make_camera(eye=(0.1, -0.2, 0.3), forward=(0.0, 0.0, -1.0)))
    got = pt.pixels()
    assert np.all(got["radiance"] == np.array([2.0, 1.0, 0.5], np.float32))
    assert np.all(got["alpha"] == 1.0) and np.all(got["object_id"] == 0)


def test_pt_back_faces_culled_in_air(ctx):
    """the same sphere, one-sided: from inside every ray sees back faces, which are culled in air"""
    m = vk.material(color=(0.0, 0.0, 0.0), emission=(1.0, 1.0, 1.0), emission_strength=2.0, two_sided=False)
    pt = vk.PathTracer(ctx, settings(40, 30, 1), seed=3)
    pt.set_scene(sphere_scene(m))
    pt.render(vk.make_camera(eye=(0.0, 0.0, 0.0), forward=(0.0, 0.0, -1.0)))
    got = pt.pixels()
    assert np.all(got["object_id"] == 0xFFFF) and np.all(got["alpha"] == 0.0)


def test_pt_batches_running_mean(ctx, scene3):
    """three batches: the stored radiance is the running mean (new + b old) / (b + 1) over batches whose
    seeds are the successive std::default_random_engine(seed) outputs"""
    w, h = 96, 54
    pt = vk.PathTracer(ctx, settings(w, h, 1, batches=8), seed=5)
    pt.set_scene(scene3)
    cam = W.room_camera()
    seeds = W.minstd_rand0(5, 3)
    acc = None
    for b, sd in enumerate(seeds):
        assert pt.current_batch() == b
        pt.render(cam)
        img = orc.pt_render(scene3, cam, w, h, 1, 8, sd)[..., :3]
        acc = img if acc is None else (img + np.float32(b) * acc) / np.float32(b + 1)
    assert pt.current_batch() == 3
    got = pt.pixels()["radiance"]
    close = np.mean(np.abs(got - acc).max(axis=2) <= 1e-4 * np.maximum(1.0, acc.max(axis=2)))
    assert close > 0.97


def test_pt_batch_state_machine(ctx):
    """max_num_batches / suspend_trace_when_done / reset_accumulator / suppress_reset / resize"""
    pt = vk.PathTracer(ctx, settings(16, 8, 1, batches=2, suspend=1), seed=1)
    pt.set_scene(empty_scene())
    cam = vk.make_camera(eye=(0.0, 0.0, 0.0), forward=(0.0, 0.0, -1.0))
    pt.render(cam)
    assert pt.current_batch() == 1 and pt.stats().traced == 1
    pt.render(cam)
    assert pt.current_batch() == 2
    pt.render(cam)                                   # done: suspended, nothing traced
    assert pt.current_batch() == 2 and pt.stats().traced == 0
    s = pt.settings
    s.suspend_trace_when_done = 0
    pt.settings = s
    pt.render(cam)                                   # keeps tracing, batch index stays capped
    assert pt.current_batch() == 2 and pt.stats().traced == 1
    pt.reset_accumulator()
    assert pt.current_batch() == 0
    s.suppress_reset = 1
    s.max_num_batches = 0                            # default: the index stays 0, every frame overwrites
    pt.settings = s
    pt.render(cam)
    assert pt.current_batch() == 0
    s.max_num_batches = 4
    pt.settings = s
    pt.render(cam)
    pt.reset_accumulator()                           # suppressed
    assert pt.current_batch() == 1
    s.resolution[0] = 20                             # new storage resets the batch index
    pt.settings = s
    pt.render(cam)
    assert pt.current_batch() == 1 and pt.pixels().shape == (8, 20)


def test_pt_deterministic(ctx, scene3):
    a = vk.PathTracer(ctx, settings(64, 36, 2), seed=11)
    b = vk.PathTracer(ctx, settings(64, 36, 2), seed=11)
    for pt in (a, b):
        pt.set_scene(scene3)
        pt.render(W.room_camera())
    assert a.pixels().tobytes() == b.pixels().tobytes()
    assert a.stats().segments > 64 * 36 * 2


def test_pt_cache_inactive_below_its_depth_is_plain_pt(ctx, scene3):
    """with the query depth at max_trace_depth no path reaches the cache: recording training vertices and
    training must leave every pixel bit-identical to plain path tracing"""
    w, h = 96, 54
    plain = vk.PathTracer(ctx, settings(w, h, 2), seed=4)
    cached = vk.PathTracer(ctx, settings(w, h, 2), seed=4)
    cache = vk.RadianceCache(ctx, vk.default_rc_params(log2_table=16), seed=1)
    for pt in (plain, cached):
        pt.set_scene(scene3)
    cached.set_radiance_cache(cache, cache_depth=8, train_stride=4)
    for pt in (plain, cached):
        pt.render(W.room_camera())
    assert plain.pixels().tobytes() == cached.pixels().tobytes()
    assert cache.step() == 2                        # one training step per sample
    cache.close()


def test_pt_cache_converges_near_reference(ctx, scene3):
    """terminating into the trained cache after two bounces: the accumulated image's mean radiance stays
    within a few percent of plain path tracing with many samples (the cache's bias), and is finite"""
    w, h = 128, 72
    ref = vk.PathTracer(ctx, settings(w, h, 256), seed=8)
    ref.set_scene(scene3)
    ref.render(W.room_camera())
    want = ref.pixels()["radiance"]
    pt = vk.PathTracer(ctx, settings(w, h, 1, batches=256), seed=9)
    pt.set_scene(scene3)
    cache = vk.RadianceCache(ctx, seed=3)
    pt.set_radiance_cache(cache, cache_depth=2, train_stride=8)
    for _ in range(192):
        pt.render(W.room_camera())
    got = pt.pixels()["radiance"]
    assert np.isfinite(got).all() and (got >= 0).all()
    rel = abs(float(got.mean()) - float(want.mean())) / float(want.mean())
    assert rel < 0.05, rel
    cache.close()


def test_transmittance_matches_oracle(ctx, scene3):
    """transmittance_test through the room scene (10% of the spheres transmissive): every record identical to
    the brute-force oracle, with and without back-face culling"""
    import oracle_binding as orc2
    rng = np.random.default_rng(21)
    n = 1500
    o = np.column_stack([rng.uniform(-2.8, 2.8, n), rng.uniform(0.2, 5.8, n), rng.uniform(-2.8, 2.8, n)])
    p = np.column_stack([rng.uniform(-2.8, 2.8, n), rng.uniform(0.2, 5.8, n), rng.uniform(-2.8, 2.8, n)])
    d = p - o
    dist = np.linalg.norm(d, axis=1)
    rays = W.make_rays(o.astype(np.float32), W.glm_normalize((d / dist[:, None]).astype(np.float32)), tmin=1e-4,
                       tmax=dist.astype(np.float32))
    pt = vk.PathTracer(ctx, settings(8, 8, 1), seed=0)
    pt.set_scene(scene3)
    for cull in (False, True):
        got = pt.trace_transmittance(rays, cull_back=cull)
        want = orc2.transmittance(scene3, rays, cull=cull)[:, :3]
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"cull={cull}"
        assert (got == 0).any(1).sum() > 100 and ((got > 0) & (got < 1)).all(1).sum() > 10
    pt.close()


@pytest.mark.parametrize("coplanar", [False, True])
def test_transmittance_many_layers_in_order(ctx, coplanar):
    """20 transmissive plates along the ray: more than one 8-candidate pass, product in ascending (t, triangle);
    coplanar: all 20 at one height, so every candidate ties in t, the passes' running bound sits exactly on
    the plates and the order comes from the triangle index alone"""
    import oracle_binding as orc2
    from test_oracle import _quad_scene, _down_rays
    zs = [0.5] * 20 if coplanar else [0.1 * k for k in range(20)]
    rng = np.random.default_rng(5)
    sc = _quad_scene(zs, list(rng.uniform(0.6, 1.0, 20)), (0.9, 0.7, 0.95))
    mats = sc["materials"].copy()
    mats["color"][:, :3] = rng.uniform(0.5, 1.0, (20, 3))
    sc["materials"] = mats
    pt = vk.PathTracer(ctx, settings(8, 8, 1), seed=0)
    pt.set_scene(sc)
    rays = _down_rays(64)
    got = pt.trace_transmittance(rays)
    want = orc2.transmittance(sc, rays)[:, :3]
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert (got > 0).all() and (got < 0.05).all()
    pt.close()


def test_denoiser_matches_oracle(ctx, scene3):
    """the a-trous denoiser over a rendered pixel buffer equals the C restatement up to exp() ulps (ocml vs
    glibc): at most 1e-4 relative per channel; raw mode is bit-exact; the tracer's image() is the same pass"""
    import oracle_binding as orc2
    w, h = 96, 54
    s = settings(w, h, 4)
    s.denoising, s.tonemap = 1, 0
    pt = vk.PathTracer(ctx, s, seed=2)
    pt.set_scene(scene3)
    pt.render(W.room_camera())
    px = pt.pixels()
    got = ctx.denoise(px)
    want = orc2.denoise(px)
    scale = np.maximum(np.abs(want[..., :3]), 1e-3)
    assert (np.abs(got[..., :3] - want[..., :3]) / scale).max() < 1e-4
    assert np.array_equal(got[..., 3], want[..., 3])
    assert np.array_equal(pt.image(), got)
    raw = ctx.denoise(px, denoise=False)
    assert np.array_equal(raw, orc2.denoise(px, denoise=False))
    # denoising lowers the display-range error against 1024 samples (values clamped to [0, 1]: the filter
    # keeps isolated fireflies and emitter pixels by design, so unclamped RMSE is theirs; measured 0.72x)
    ref = vk.PathTracer(ctx, settings(w, h, 1024), seed=3)
    ref.set_scene(scene3)
    ref.render(W.room_camera())
    r = np.clip(ref.pixels()["radiance"], 0, 1)
    rmse = lambda x: np.sqrt(np.mean((np.clip(x, 0, 1) - r) ** 2))
    assert rmse(got[..., :3]) < 0.8 * rmse(px["radiance"])
    pt.close()
    ref.close()


def _within_half_ulp(got, want):
    """equal, or one binary16 step apart (expf of ocml and glibc differ in the last float32 ulp, which can
    move a value across a rounding boundary of the half-precision target)"""
    step = np.spacing(np.maximum(np.abs(got), np.abs(want)).astype(np.float16)).astype(np.float32)
    return (np.abs(got - want) <= step).all() and np.mean(got == want) > 0.99


def test_tonemap_matches_oracle(ctx, scene3):
    """the composited image (bloom_composition.slang without bloom) of the denoised and of the raw image
    against the C restatement; the tracer's image() is the same fused pass as vkrt_tonemap after
    vkrt_denoise, bit for bit"""
    import oracle_binding as orc2
    w, h = 80, 45
    for denoising, gamma, exposure in ((1, 1.0, 2.0), (0, 2.2, 1.0), (1, 2.2, 0.5)):
        s = settings(w, h, 2)
        s.denoising, s.tonemap, s.gamma, s.exposure = denoising, 1, gamma, exposure
        pt = vk.PathTracer(ctx, s, seed=4)
        pt.set_scene(scene3)
        pt.render(W.room_camera())
        px = pt.pixels()
        img = pt.image()
        lin = ctx.denoise(px, denoise=bool(denoising))
        assert np.array_equal(img, ctx.tonemap(lin, exposure, gamma))
        want = orc2.tonemap(orc2.denoise(px, denoise=bool(denoising)), exposure, gamma)
        assert _within_half_ulp(img, want), (denoising, gamma, exposure)
        assert np.array_equal(img, img.astype(np.float16).astype(np.float32))
        assert (img[..., :3] >= 0).all() and (img[..., :3] <= 1).all()
        pt.close()
    rng = np.random.default_rng(6)
    tex = np.concatenate([rng.exponential(1.0, (5000, 4)), np.zeros((1, 4)), np.full((1, 4), 1e4)]).astype(np.float32)
    for gamma, exposure in ((1.0, 2.0), (2.2, 1.0), (0.7, 3.5)):
        assert _within_half_ulp(ctx.tonemap(tex, exposure, gamma), orc2.tonemap(tex, exposure, gamma))
    with pytest.raises(vk.VkrtError):
        ctx.tonemap(tex, 2.0, 0.0)
