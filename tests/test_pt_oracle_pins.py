"""Pins the CPU reference integrator's BSDF sampling, Russian roulette and denoiser against mathematics that
does not come from its own code: chi-square tests of sampled directions against densities (eval_disney for
the lobes whose sampler and pdf agree, and the published closed forms, Heitz's visible-normal density and
Burley's GTR1, for the two lobes where the reference's sampler does not draw from its returned pdf), a mirror
furnace and an index-matched glass sphere with closed-form images, Russian roulette's unbiasedness, and
a-trous passes worked out by hand from shaders/slang/denoise.slang. Reference lines:
bsdf_disney.slang:162-296 (sample_disney), :298-381 (eval_disney), raypipeline.slang:667-677 (roulette)."""
import numpy as np
import pytest
from scipy import stats

import oracle_binding as orc
import vierkant_amd as vk
from vierkant_amd import workloads as W

N = np.array([0.0, 0.0, 1.0], np.float32)
V = np.array([np.sin(0.7) * np.cos(0.5), np.sin(0.7) * np.sin(0.5), np.cos(0.7)], np.float32)
V /= np.linalg.norm(V)


def mat(**kw):
    return np.array([vk.material(**kw)], vk.MATERIAL_DTYPE)


def rng_draws(seed, count, k):
    """the first k rnd() draws of every sample's RNG chain: state xxhash32(seed, i) (utils/hash.slang:9-18,
    vectorized here and checked against the oracle's, which test_oracle pins to the xxhash package), then
    lcg steps with rnd = (state & 0xFFFFFF) / 2^24 (utils/random.slang:27-36)"""
    with np.errstate(over="ignore"):
        u = np.uint32
        i = np.arange(count, dtype=np.uint32)
        h = u(seed) + u(374761393) + i * u(3266489917)
        h = u(668265263) * ((h << u(17)) | (h >> u(15)))
        h = u(2246822519) * (h ^ (h >> u(15)))
        h = u(3266489917) * (h ^ (h >> u(13)))
        h = h ^ (h >> u(16))
        for j in (0, 1, count // 2, count - 1):
            assert int(h[j]) == orc.xxhash32(seed, j)
        out = np.zeros((count, k))
        for j in range(k):
            h = u(1664525) * h + u(1013904223)
            out[:, j] = (h & u(0xFFFFFF)).astype(np.float64) / 2.0 ** 24
    first, _ = orc.rnd_sequence(orc.xxhash32(seed, 1), k)
    assert np.allclose(out[1], first, rtol=0, atol=0)
    return out


def grid(n_ct, n_phi, lo=0.0, sub=8):
    """cells in (cos theta, phi) about +z and, per cell, sub x sub midpoint directions with their solid angle"""
    ct_e = np.linspace(lo, 1.0, n_ct + 1)
    ph_e = np.linspace(0.0, 2 * np.pi, n_phi + 1)
    u = (np.arange(sub) + 0.5) / sub
    ct = (ct_e[:-1, None] + np.diff(ct_e)[:, None] * u[None, :]).reshape(-1)
    ph = (ph_e[:-1, None] + np.diff(ph_e)[:, None] * u[None, :]).reshape(-1)
    C, P = np.meshgrid(ct, ph, indexing="ij")
    st = np.sqrt(np.maximum(0.0, 1 - C * C))
    dirs = np.stack([st * np.cos(P), st * np.sin(P), C], -1).reshape(-1, 3)
    d_omega = (ct_e[1] - ct_e[0]) / sub * (2 * np.pi / n_phi) / sub
    return ct_e, ph_e, dirs, d_omega, sub


def cell_integral(density, n_ct, n_phi, sub):
    return density.reshape(n_ct, sub, n_phi, sub).sum(axis=(1, 3))


def chi2_pvalue(obs, exp, min_expected=10.0):
    """Pearson chi-square over cells, those expecting fewer than min_expected counts pooled into one"""
    obs, exp = np.asarray(obs, np.float64).ravel(), np.asarray(exp, np.float64).ravel()
    small = exp < min_expected
    o = np.concatenate([obs[~small], [obs[small].sum()]])
    e = np.concatenate([exp[~small], [exp[small].sum()]])
    keep = e > 0
    stat = np.sum((o[keep] - e[keep]) ** 2 / e[keep])
    return stats.chi2.sf(stat, keep.sum() - 1), stat


def binned(dirs, ct_e, ph_e):
    phi = np.mod(np.arctan2(dirs[:, 1], dirs[:, 0]), 2 * np.pi)
    h, _, _ = np.histogram2d(dirs[:, 2], phi, bins=[ct_e, ph_e])
    return h


def eval_density(material, eta, dirs):
    return orc.eval_disney(material, N, V, eta, dirs)[:, 3].astype(np.float64)


@pytest.mark.parametrize("name,kw,sphere", [
    ("specular (metal, GGX VNDF)", dict(metalness=1.0, roughness=0.6, color=(0.9, 0.8, 0.7)), False),
    ("diffuse + specular (dielectric base)", dict(metalness=0.0, roughness=0.7, color=(0.8, 0.8, 0.8)), False),
    ("diffuse transmission", dict(metalness=0.0, roughness=0.8, diffuse_transmission=1.0,
                                  diffuse_transmission_color=(0.6, 0.7, 0.8, 0.0)), True),
])
def test_sampled_directions_follow_eval_pdf(name, kw, sphere):
    """chi-square of 300,000 sampled directions against eval_disney's pdf integrated over each cell (the
    lobe mixture's density); samples the sampler returns with pdf 0 (below the horizon) form one more cell,
    whose expectation is the density's missing mass"""
    m = mat(**kw)
    count = 300000
    s = orc.sample_disney(m, N, V, 1.0, 11, count)
    valid = s[:, 3] > 0
    n_ct, n_phi = (24, 16) if sphere else (12, 16)
    ct_e, ph_e, dirs, dw, sub = grid(n_ct, n_phi, -1.0 if sphere else 0.0)
    dens = eval_density(m, 1.0, dirs)
    exp = cell_integral(dens, n_ct, n_phi, sub) * dw * count
    obs = binned(s[valid, :3], ct_e, ph_e)
    p, stat = chi2_pvalue(np.append(obs, count - valid.sum()), np.append(exp, count - exp.sum()))
    assert p > 1e-4, f"{name}: chi2 {stat:.1f}, p {p:.2e}"
    # and the test has power: the same samples against a cosine density fail it
    cosine = np.maximum(dirs[:, 2], 0) / np.pi if not sphere else np.abs(dirs[:, 2]) / (2 * np.pi)
    wrong = cell_integral(cosine, n_ct, n_phi, sub) * dw * count
    assert chi2_pvalue(np.append(obs, count - valid.sum()), np.append(wrong, count - wrong.sum()))[0] < 1e-6


def test_specular_lobe_returns_its_pdf():
    """single-lobe material: every draw's returned pdf is eval_disney's pdf at the drawn direction"""
    m = mat(metalness=1.0, roughness=0.4)
    s = orc.sample_disney(m, N, V, 1.0, 5, 20000)
    ok = s[:, 3] > 0
    ev = orc.eval_disney(m, N, V, 1.0, s[ok, :3])
    assert np.allclose(ev[:, 3], s[ok, 3], rtol=2e-4)
    assert np.allclose(ev[:, :3], s[ok, 4:7], rtol=2e-4)


def vndf_density(h, alpha):
    """Heitz 2018: D_V(h) = G1(V) max(0, V.h) D(h) / (N.V) for isotropic GGX, as a density of h"""
    nh = h[:, 2]
    a2 = alpha * alpha
    d = a2 / (np.pi * (1 + (a2 - 1) * nh * nh) ** 2)
    nv = float(V[2])
    g1 = 2 * nv / (nv + np.sqrt(a2 + (1 - a2) * nv * nv))
    return g1 * np.maximum(0.0, h @ V.astype(np.float64)) * d * (nh > 0) / nv


def fresnel_dielectric(ci, eta):
    s2 = eta * eta * (1 - ci * ci)
    ct = np.sqrt(np.maximum(1 - s2, 0))
    rs = (eta * ct - ci) / (eta * ct + ci)
    rp = (eta * ci - ct) / (eta * ci + ct)
    return np.where(s2 >= 1, 1.0, 0.5 * (rs * rs + rp * rp))


@pytest.mark.parametrize("eta", [1 / 1.5, 1.5])
def test_dielectric_half_vectors_and_branch_rule(eta):
    """transmission 1: the half vectors recovered from the drawn directions (L + V for reflections,
    L + eta V for refractions) follow Heitz's visible-normal density, and the reflect/refract choice is the
    reference's: reflect iff the VNDF sample's own second draw Xi.y < F(|V.h|) (bsdf_disney.slang:188, the
    draw is reused), or on total internal reflection. Because the draw is reused, reflections come from the
    half vectors nearest N and the returned pdf (F times the VNDF density) is not the density of the drawn
    reflections; this pins that behaviour instead of an ideal Fresnel split."""
    rough = 0.5
    alpha = max(0.001, rough * rough)
    m = mat(metalness=0.0, roughness=rough, transmission=1.0, ior=1.5)
    count = 200000
    s = orc.sample_disney(m, N, V, eta, 21, count)
    L = s[:, :3].astype(np.float64)
    refr = s[:, 7] > 0
    H = np.where(refr[:, None], L + eta * V, L + V)
    H /= np.linalg.norm(H, axis=1, keepdims=True)
    H *= np.sign(H @ V.astype(np.float64))[:, None]
    ct_e, ph_e, dirs, dw, sub = grid(16, 16, sub=48)
    dens = vndf_density(dirs, alpha)
    assert abs(dens.sum() * dw - 1.0) < 2e-4  # the quadrature resolves the peak
    exp = cell_integral(dens, 16, 16, sub) * dw * count
    obs = binned(H, ct_e, ph_e)
    p, stat = chi2_pvalue(np.append(obs, count - obs.sum()), np.append(exp, count - exp.sum()))
    assert p > 1e-4, f"eta {eta}: chi2 {stat:.1f}, p {p:.2e}"
    xi_y = rng_draws(21, count, 2)[:, 1]
    F = fresnel_dielectric(np.abs(H @ V.astype(np.float64)), eta)
    tir = eta * eta * (1 - (H @ V.astype(np.float64)) ** 2) >= 1
    want_refl = (xi_y < F) | tir
    near = np.abs(xi_y - F) < 1e-5
    assert np.all((~refr == want_refl) | near), np.sum((~refr != want_refl) & ~near)
    assert 0.001 < np.mean(~refr) < 0.999  # both branches occur (at eta 1.5 mostly by total reflection)


def gtr1_density(h, alpha):
    """Burley 2012: D_GTR1(h) = (a^2 - 1) / (pi ln a^2 (1 + (a^2 - 1) cos^2)), density of h = D cos"""
    a2 = alpha * alpha
    c = h[:, 2]
    return (a2 - 1) / (np.pi * np.log(a2) * (1 + (a2 - 1) * c * c)) * c * (c > 0)


def test_clearcoat_half_vectors_follow_gtr1_and_report_gtr2():
    """clearcoat 0.5 on a dielectric base: the draws the RNG chain sends to the clearcoat lobe (draw 4 >= the
    diffuse ratio 0.5, draw 5 >= the primary-specular ratio 1 / 1.5) have half vectors distributed as GTR1
    with alpha = lerp(0.1, 0.001, clearcoat) (the sampler's roughness), while the returned pdf is
    (1 - 1/1.5) 0.5 GTR2(N.h, clearcoat_roughness) N.h / (4 V.h) (EvalClearcoat's roughness): the two differ
    in the reference (bsdf_disney.slang:279-286), and this pins both"""
    cc, ccr = 0.5, 0.3
    m = mat(metalness=0.0, roughness=0.5, clearcoat=cc, clearcoat_roughness=ccr)
    count = 400000
    s = orc.sample_disney(m, N, V, 1.0, 31, count)
    r = rng_draws(31, count, 5)
    sel = (r[:, 3] >= 0.5) & (r[:, 4] >= 1 / (1 + cc))
    L = s[sel, :3].astype(np.float64)
    H = L + V
    H /= np.linalg.norm(H, axis=1, keepdims=True)
    H *= np.where(H[:, 2] < 0, -1.0, 1.0)[:, None]
    alpha_s = 0.1 + (0.001 - 0.1) * cc
    a2 = alpha_s * alpha_s
    # cells of equal probability under GTR1: integrating the density over the cap cos theta <= c gives
    # P = ln(1 + (a^2 - 1) c^2) / ln a^2, so the k-th of K edges is c = sqrt((a^(2k/K) - 1) / (a^2 - 1))
    K, n_phi = 20, 8
    ct_e = np.sqrt((a2 ** (np.arange(K + 1) / K) - 1) / (a2 - 1))
    ct_e[0], ct_e[-1] = 0.0, 1.0 + 1e-12
    obs = binned(H, ct_e, np.linspace(0, 2 * np.pi, n_phi + 1))
    p, stat = chi2_pvalue(obs, np.full(obs.shape, len(H) / (K * n_phi)))
    assert p > 1e-4, f"chi2 {stat:.1f}, p {p:.2e}"
    # the density form agrees with that CDF (a numeric check of the calculus, not of the sampler)
    c = np.linspace(0, 1, 200001)
    pdf_c = 2 * np.pi * gtr1_density(np.stack([np.sqrt(1 - c * c), 0 * c, c], 1), alpha_s)
    cdf = np.concatenate([[0], np.cumsum(0.5 * (pdf_c[1:] + pdf_c[:-1]) * np.diff(c))])
    assert np.allclose(cdf, np.log1p((a2 - 1) * c * c) / np.log(a2), atol=2e-4)
    # the returned pdf: GTR2 at the clearcoat roughness, zero below the horizon
    nh, vh = H[:, 2], H @ V.astype(np.float64)
    a2 = ccr * ccr
    gtr2 = a2 / (np.pi * (1 + (a2 - 1) * nh * nh) ** 2)
    want = (1 - 1 / (1 + cc)) * 0.5 * gtr2 * nh / (4 * vh)
    up = L[:, 2] > 0
    assert np.allclose(s[sel, 3][up], want[up], rtol=1e-4)
    assert np.all(s[sel, 3][~up] == 0)


# ---- closed-form images ---------------------------------------------------------------------------------

def quad_scene(material, size=50.0):
    pos = np.array([[-size, 0, -size], [size, 0, -size], [size, 0, size], [-size, 0, size]], np.float32)
    return {"positions": pos, "normals": np.tile(np.float32([0, 1, 0]), (4, 1)),
            "indices": np.array([0, 2, 1, 0, 3, 2], np.uint32), "tri_object": np.zeros(2, np.uint32),
            "object_material": np.zeros(1, np.uint32), "materials": np.array([material], vk.MATERIAL_DTYPE),
            "environment": 1.0}


def sky(d):
    d = np.asarray(d, np.float64)
    return np.where(d[..., 1] > 0, 0.6 + (1.3 - 0.6) * d[..., 1], 0.6 + (0.1 - 0.6) * (-d[..., 1]))


def camera_dirs(cam, w, h):
    """pixel-center directions of the pinhole camera (the sky-only test's mapping, offset 0)"""
    eye, fw, rt, up, thx, thy = [np.asarray(v, np.float64) for v in W.camera_basis(cam, w, h)]
    y, x = np.mgrid[0:h, 0:w]
    u, v = (x + 0.5) / w, (y + 0.5) / h
    d = rt * ((2 * u - 1) * thx)[..., None] + up * (-(2 * v - 1) * thy)[..., None] + fw
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def test_white_mirror_furnace():
    """a white metal (color 1, metalness 1, roughness 0.03: alpha = 0.001) plane under environment_white:
    F = 1 and G2/G1 ~ 1, so every pixel that sees the plane returns the sky in the mirrored direction; with a
    sample offset per pixel the camera ray varies by up to a pixel, covered by the tolerance"""
    m = vk.material(color=(1.0, 1.0, 1.0), metalness=1.0, roughness=0.03)
    w, h = 48, 32
    cam = vk.make_camera(eye=(0.0, 2.0, 0.0), forward=(0.0, -0.6, -1.0), fov_y=0.6)
    img = orc.pt_render(quad_scene(m), cam, w, h, 16, 4, 5)
    d = camera_dirs(cam, w, h)
    hit = img[..., 3] > 0
    want = sky(d * np.array([1.0, -1.0, 1.0]))[hit]
    got = img[..., 0][hit]
    assert hit.sum() > 0.9 * w * h and np.all(img[..., 0] == img[..., 2])
    # energy: the mean over the plane is the mirrored sky's (GGX's heavy tail scatters single samples)
    assert abs(got.mean() / want.mean() - 1.0) < 1e-3
    assert np.quantile(np.abs(got - want), 0.99) < 0.01


def test_index_matched_glass_sphere_is_invisible():
    """transmission 1, ior 1, white, nearly smooth: eta = 1 gives F = 0 and refract(-V, h, 1) = -V for
    every half vector, so paths pass straight through both surfaces with weight ~1 and each pixel returns
    the sky behind it, as the image without the sphere does"""
    pos, idx = vk.icosphere(1.0, 3)
    nrm = (pos / np.linalg.norm(pos, axis=1, keepdims=True)).astype(np.float32)
    m = vk.material(color=(1.0, 1.0, 1.0), transmission=1.0, ior=1.0, roughness=0.03)
    sc = {"positions": pos, "normals": nrm, "indices": idx, "tri_object": np.zeros(len(idx) // 3, np.uint32),
          "object_material": np.zeros(1, np.uint32), "materials": np.array([m], vk.MATERIAL_DTYPE),
          "environment": 1.0}
    empty = dict(sc, positions=np.zeros((0, 3), np.float32), normals=np.zeros((0, 3), np.float32),
                 indices=np.zeros(0, np.uint32), tri_object=np.zeros(0, np.uint32))
    cam = vk.make_camera(eye=(0.0, 0.2, 3.0), forward=(0.0, -0.05, -1.0), fov_y=0.9)
    w, h = 40, 30
    img = orc.pt_render(sc, cam, w, h, 4, 6, 3)
    ref = orc.pt_render(empty, cam, w, h, 4, 6, 3)
    seen = img[..., 3] > 0
    assert seen.sum() > 0.25 * w * h
    assert np.allclose(img[..., :3], ref[..., :3], rtol=2e-3, atol=1e-4)


def test_russian_roulette_is_unbiased():
    """the room at 64 x 36, 8 spp, depth 8, in 8 independent seeds with the reference's roulette
    (raypipeline.slang:667-677) and without it: the images differ (roulette acts) while the image means agree
    within 4 standard errors of their difference"""
    sc = W.room_scene(n_spheres=20)
    cam = W.room_camera()
    with_rr, without = [], []
    for seed in range(8):
        a = orc.pt_render(sc, cam, 64, 36, 8, 8, 100 + seed)
        b = orc.pt_render(sc, cam, 64, 36, 8, 8, 100 + seed, test_flags=orc.PT_NO_ROULETTE)
        assert not np.array_equal(a[..., :3], b[..., :3])
        with_rr.append(float(a[..., :3].mean()))
        without.append(float(b[..., :3].mean()))
    d = np.array(with_rr) - np.array(without)
    se = d.std(ddof=1) / np.sqrt(len(d))
    assert abs(d.mean()) < 4 * se + 1e-6, (d.mean(), se)


# ---- a-trous denoiser by hand (shaders/slang/denoise.slang) ----------------------------------------------

KERNEL_1D = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])  # the 5x5 table is its outer product


def test_atrous_first_pass_by_hand():
    """5 x 5 image, equal normals, depths and albedo (0.5): a bright centre pixel (radiance 1.0, luminance
    2.0 after demodulation) in a field of 0.25 (luminance 0.5). Colour weight exp(-(dl)^2 / 0.9) with the
    NTSC luminance, taps clamped to the edge, output = sum(w c / a) / sum(w) * a"""
    from test_pt_oracle import _pixels
    rad = np.full((5, 5, 3), 0.25, np.float32)
    rad[2, 2] = 1.0
    px = _pixels(5, 5, rad, albedo=(0.5, 0.5, 0.5))
    out = orc.denoise_passes(px, 1)
    a = 127 / 255.0  # the albedo 0.5 packed as 8 bits (int(0.5 * 255) = 127) and unpacked
    dm = rad[..., 0].astype(np.float64) / a
    lum = dm  # grey: the NTSC weights sum to 1
    cphi = 1 / 0.9
    for (y, x) in [(2, 2), (0, 0), (1, 2), (0, 4)]:
        num = den = 0.0
        for j in range(5):
            for i in range(5):
                yy, xx = min(max(y + j - 2, 0), 4), min(max(x + i - 2, 0), 4)
                wgt = KERNEL_1D[j] * KERNEL_1D[i] * min(np.exp(-(lum[y, x] - lum[yy, xx]) ** 2 * cphi), 1.0)
                num += wgt * dm[yy, xx]
                den += wgt
        want = num / den * a
        assert np.allclose(out[y, x, :3], want, rtol=2e-6), ((y, x), out[y, x], want)


def test_atrous_second_pass_uses_step_two_and_the_first_pass_output():
    """the second pass reads the first pass's image and taps at offsets 2 (-2, -1, 0, 1, 2), clamped"""
    from test_pt_oracle import _pixels
    rng = np.random.default_rng(5)
    rad = rng.uniform(0.1, 1.0, (7, 6, 3)).astype(np.float32)
    px = _pixels(7, 6, rad, albedo=(1.0, 1.0, 1.0))
    first = orc.denoise_passes(px, 1)[..., :3].astype(np.float64)
    second = orc.denoise_passes(px, 2)[..., :3]
    lw = np.array([0.298999994993, 0.587000012398, 0.114000000060])
    lum = first @ lw
    h, w = lum.shape
    for (y, x) in [(3, 3), (0, 0), (6, 5), (1, 4)]:
        num, den = np.zeros(3), 0.0
        for j in range(5):
            for i in range(5):
                yy, xx = min(max(y + 2 * (j - 2), 0), h - 1), min(max(x + 2 * (i - 2), 0), w - 1)
                wgt = KERNEL_1D[j] * KERNEL_1D[i] * min(np.exp(-(lum[y, x] - lum[yy, xx]) ** 2 / 0.9), 1.0)
                num += wgt * first[yy, xx]
                den += wgt
        assert np.allclose(second[y, x], num / den, rtol=2e-5), ((y, x), second[y, x], num / den)


# ---- how far one ulp of the math library moves the room image --------------------------------------------

ROOM_W, ROOM_H, ROOM_SPP = 160, 90, 4
# the share of pixels within 1e-4 and the RMSE (relative to the mean) between the room image and the same
# image with every transcendental one float up; the GPU parity test holds its tracer to these
ULP_SHARE_FLOOR = 0.97
ULP_RMSE_CEIL = 0.02


def test_one_ulp_sensitivity_of_the_room_image():
    """the room test image of the GPU parity test, rendered twice by the integrator: as is, and with every
    cos/sin/tan/atan/asin/pow/log2/exp result moved one float up. The share of pixels that stay within 1e-4
    and the RMSE between the two bound what two correct implementations with different math libraries can
    reach; the GPU test's thresholds are these measured values"""
    sc = W.room_scene()
    cam = W.room_camera()
    seed = W.minstd_rand0(0, 1)[0]
    a = orc.pt_render(sc, cam, ROOM_W, ROOM_H, ROOM_SPP, 8, seed)[..., :3]
    b = orc.pt_render(sc, cam, ROOM_W, ROOM_H, ROOM_SPP, 8, seed, test_flags=orc.PT_ULP_BIAS)[..., :3]
    diff = np.abs(a - b).max(axis=2)
    share = float(np.mean(diff <= 1e-4 * np.maximum(1.0, a.max(axis=2))))
    rmse = float(np.sqrt(np.mean((a - b) ** 2))) / float(a.mean())
    print(f"one-ulp sensitivity: {share:.4f} of pixels within 1e-4, relative rmse {rmse:.5f}")
    assert share < 1.0 and rmse > 0  # the perturbation reaches the image
    assert share >= ULP_SHARE_FLOOR and rmse <= ULP_RMSE_CEIL
