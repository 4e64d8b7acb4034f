"""Pins the textured-material parts of the CPU integrator (oracle/vkpt_oracle.c) and the library's image
decoding: the mip chain and SampleLevel against scipy's periodic linear interpolation, PNG decoding against
Pillow, the alpha-mask any-hit stages against renders that differ only by the masked surface, and the ray
cone's level of detail against geometry worked out here (raypipeline.slang:280-321, 526-531, 608, 686-720;
visibility_test.slang:136-155)."""
import struct
import zlib

import numpy as np
import pytest
from scipy import ndimage

import oracle_binding as orc
import vierkant_amd as vk
from vierkant_amd import workloads as W


# ---- mip chain and sampling ---------------------------------------------------------------------------------
def mip_chain(img):
    """levels of an (h, w, 4) uint8 image: 2 x 2 box filter with edge clamping, (a + b + c + d + 2) // 4"""
    levels = [img.astype(np.int64)]
    while levels[-1].shape[0] > 1 or levels[-1].shape[1] > 1:
        p = levels[-1]
        h, w = p.shape[:2]
        nh, nw = max(1, h // 2), max(1, w // 2)
        ys = np.minimum(np.stack([2 * np.arange(nh), 2 * np.arange(nh) + 1]), h - 1)
        xs = np.minimum(np.stack([2 * np.arange(nw), 2 * np.arange(nw) + 1]), w - 1)
        acc = sum(p[ys[a]][:, xs[b]] for a in range(2) for b in range(2))
        levels.append((acc + 2) // 4)
    return [lv.astype(np.float64) / 255.0 for lv in levels]


def sample_level(level, u, v):
    """bilinear filtering with repeat addressing, texel centres at (i + 0.5) / size (scipy grid-wrap)"""
    h, w = level.shape[:2]
    coords = [v * h - 0.5, u * w - 0.5]
    return np.stack([ndimage.map_coordinates(level[..., c], coords, order=1, mode="grid-wrap") for c in range(4)], -1)


def sample(levels, u, v, lam):
    lam = np.nan_to_num(np.asarray(lam, np.float64), nan=0.0, neginf=0.0, posinf=len(levels) - 1)
    lam = np.clip(lam, 0, len(levels) - 1)
    l0 = np.floor(lam).astype(int)
    f = (lam - l0)[:, None]
    out = np.zeros((len(u), 4))
    for k in range(len(levels)):
        sel = l0 == k
        if sel.any():
            a = sample_level(levels[k], u[sel], v[sel])
            b = sample_level(levels[min(k + 1, len(levels) - 1)], u[sel], v[sel])
            out[sel] = a + f[sel] * (b - a)
    return out


@pytest.mark.parametrize("shape", [(16, 16), (7, 12), (1, 9), (33, 5)])
def test_sampler_matches_periodic_linear_interpolation(shape):
    """SampleLevel over the box-filtered chain, odd and one-texel sizes included: scipy's grid-wrap linear
    interpolation on each level, linear between levels, within float32 rounding"""
    rng = np.random.default_rng(sum(shape))
    img = rng.integers(0, 256, (*shape, 4), dtype=np.uint8)
    n = 4000
    uv = rng.uniform(-3.0, 3.0, (n, 2)).astype(np.float32)
    lam = rng.uniform(-1.0, 7.0, n).astype(np.float32)
    lam[:5] = [np.nan, -np.inf, np.inf, 0.0, 100.0]
    got = orc.texture_sample(img, uv, lam)
    want = sample(mip_chain(img), uv[:, 0].astype(np.float64), uv[:, 1].astype(np.float64), lam)
    assert np.abs(got - want).max() < 2e-5


def test_mip_levels_of_a_constant_and_a_checkerboard():
    """a constant image keeps its value on every level; a one-texel checkerboard of 0 and 255 averages to
    (0 + 255 + 0 + 255 + 2) // 4 = 128 from level 1 on, at any coordinate"""
    c = np.full((8, 8, 4), 77, np.uint8)
    uv = np.random.default_rng(0).uniform(0, 1, (64, 2)).astype(np.float32)
    for lam in (0.0, 1.5, 3.0):
        assert np.allclose(orc.texture_sample(c, uv, lam), 77 / 255, atol=1e-6)
    chk = np.zeros((8, 8, 4), np.uint8)
    chk[(np.arange(8)[:, None] + np.arange(8)) % 2 == 0] = 255
    for lam in (1.0, 2.0, 3.0):
        assert np.allclose(orc.texture_sample(chk, uv, lam), 128 / 255, atol=1e-6)
    # texel centres of level 0 read the texels themselves
    ij = np.stack(np.meshgrid(np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 2)
    centres = ((ij[:, ::-1] + 0.5) / 8).astype(np.float32)
    got = orc.texture_sample(chk, centres, 0.0)
    assert np.allclose(got, chk.reshape(-1, 4) / 255, atol=1e-6)


# ---- PNG decoding -------------------------------------------------------------------------------------------
def _chunk(kind, body):
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def png_bytes(rows_by_pass, width, height, depth, ctype, interlace=0, filters=(0,), plte=None, trns=None):
    """a PNG from raw scanlines (bytes per row, per pass), each row given the next filter type in turn and
    encoded by it (the decoder must undo each)"""
    out = b""
    k = 0
    for rows in rows_by_pass:
        prev = None
        for r in rows:
            f = filters[k % len(filters)]
            k += 1
            bpp = max(1, {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype] * depth // 8)
            cur = np.frombuffer(r, np.uint8).astype(np.int64)
            up = np.zeros_like(cur) if prev is None else prev
            enc = np.zeros_like(cur)
            for x in range(len(cur)):
                a = cur[x - bpp] if x >= bpp else 0
                b = up[x]
                c = up[x - bpp] if x >= bpp else 0
                if f == 0:
                    p = 0
                elif f == 1:
                    p = a
                elif f == 2:
                    p = b
                elif f == 3:
                    p = (a + b) // 2
                else:
                    pp = a + b - c
                    pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                enc[x] = (cur[x] - p) % 256
            out += bytes([f]) + enc.astype(np.uint8).tobytes()
            prev = cur
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", width, height, depth, ctype, 0, 0,
                                                                 interlace))
    if plte is not None:
        data += _chunk(b"PLTE", plte)
    if trns is not None:
        data += _chunk(b"tRNS", trns)
    return data + _chunk(b"IDAT", zlib.compress(out, 9)) + _chunk(b"IEND", b"")


def pack_rows(samples, depth):
    """(h, w*channels) integer samples -> packed scanlines at the bit depth"""
    rows = []
    for r in samples:
        if depth == 16:
            rows.append(np.asarray(r, ">u2").tobytes())
        elif depth == 8:
            rows.append(np.asarray(r, np.uint8).tobytes())
        else:
            bits = np.unpackbits(np.asarray(r, np.uint8)[:, None], axis=1)[:, 8 - depth:].reshape(-1)
            rows.append(np.packbits(bits).tobytes())
    return rows


ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def encode(samples, depth, ctype, channels, interlace, **kw):
    h, w = samples.shape[:2]
    if not interlace:
        return png_bytes([pack_rows(samples.reshape(h, w * channels), depth)], w, h, depth, ctype, **kw)
    passes = []
    for x0, y0, dx, dy in ADAM7:
        sub = samples[y0::dy, x0::dx]
        if sub.size:
            passes.append(pack_rows(sub.reshape(sub.shape[0], -1), depth))
    return png_bytes(passes, w, h, depth, ctype, interlace=1, **kw)


def load_png(tmp_path, data, name="t.png"):
    (tmp_path / name).write_bytes(data)
    (tmp_path / "m.mtl").write_text(f"newmtl m\nmap_Kd {name}\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl m\nf 1 2 3\n")
    m = vk.load_model(str(tmp_path / "m.obj"))
    assert m["n_skipped_textures"] == 0 and len(m["textures"]) == 1
    return m["textures"][0]


CASES = [  # depth, ctype, channels, interlace, filters
    (8, 6, 4, 0, (0, 1, 2, 3, 4)), (8, 2, 3, 0, (4, 3, 2, 1, 0)), (8, 0, 1, 0, (1, 4)), (8, 4, 2, 0, (3,)),
    (16, 6, 4, 0, (4, 2)), (16, 0, 1, 0, (1,)), (4, 0, 1, 0, (2, 4)), (2, 0, 1, 0, (0,)), (1, 0, 1, 0, (1, 3)),
    (8, 6, 4, 1, (0, 1, 2, 3, 4)), (16, 2, 3, 1, (4,)), (1, 0, 1, 1, (2,)),
]


@pytest.mark.parametrize("depth,ctype,channels,interlace,filters", CASES)
def test_png_decoding_matches_pillow(tmp_path, depth, ctype, channels, interlace, filters):
    """every colour type, bit depth, row filter and Adam7 interlacing, odd sizes: the RGBA8 texels equal
    Pillow's reading of the same file (16-bit samples keep their high byte; grey below 8 bits scales to
    0..255)"""
    from PIL import Image
    rng = np.random.default_rng(depth * 10 + ctype + 100 * interlace)
    h, w = 13, 11
    samples = rng.integers(0, 1 << depth, (h, w, channels))
    data = encode(samples, depth, ctype, channels, interlace, filters=filters)
    got = load_png(tmp_path, data)
    (tmp_path / "p.png").write_bytes(data)
    im = Image.open(tmp_path / "p.png")
    if depth == 16 and ctype == 0:
        # Pillow keeps 16-bit grey at full depth (mode I;16); its high byte is the 8-bit value
        g = np.asarray(im).astype(np.int64)[..., None] >> 8
        want = np.concatenate([g, g, g, np.full((h, w, 1), 255)], -1)
    else:
        want = np.asarray(im.convert("RGBA")).astype(np.int64)
    assert np.array_equal(got.astype(np.int64), want)


def test_png_palette_and_transparency_keys(tmp_path):
    """a 4-bit palette with tRNS alphas, and grey / RGB colour keys that make one value transparent"""
    from PIL import Image
    rng = np.random.default_rng(5)
    pal = rng.integers(0, 256, (16, 3), dtype=np.uint8)
    alphas = rng.integers(0, 256, 10, dtype=np.uint8)
    idx = rng.integers(0, 16, (9, 14, 1))
    data = encode(idx, 4, 3, 1, 0, filters=(1, 2), plte=pal.tobytes(), trns=alphas.tobytes())
    got = load_png(tmp_path, data)
    (tmp_path / "p.png").write_bytes(data)
    want = np.asarray(Image.open(tmp_path / "p.png").convert("RGBA"))
    assert np.array_equal(got, want)
    grey = rng.integers(0, 4, (6, 5, 1))
    data = encode(grey, 8, 0, 1, 0, trns=struct.pack(">H", 2))
    got = load_png(tmp_path, data, "g.png")
    assert np.array_equal(got[..., 3], np.where(grey[..., 0] == 2, 0, 255))
    assert np.array_equal(got[..., 0], grey[..., 0])
    rgb = rng.integers(0, 2, (6, 5, 3))
    data = encode(rgb, 8, 2, 3, 1, trns=struct.pack(">HHH", 1, 0, 1))
    got = load_png(tmp_path, data, "c.png")
    key = (rgb == [1, 0, 1]).all(-1)
    assert np.array_equal(got[..., 3], np.where(key, 0, 255))


def test_obj_coordinates_flip_only_under_a_diffuse_texture(tmp_path):
    """the diffuse texture's transform scales v by -1 (wavefront_obj.cpp:100-111); faces of other materials
    keep their coordinates"""
    Image = pytest.importorskip("PIL.Image")
    Image.fromarray(np.full((2, 2, 4), 200, np.uint8)).save(tmp_path / "t.png")
    (tmp_path / "m.mtl").write_text("newmtl tex\nmap_Kd t.png\nnewmtl plain\nKd 0.5 0.5 0.5\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.25 0.75\nvt 1 0.5\nvt 0 1\n"
                                    "o a\nusemtl tex\nf 1/1 2/2 3/3\no b\nusemtl plain\nf 1/1 2/2 3/3\n")
    m = vk.load_model(str(tmp_path / "m.obj"))
    want = np.float32([[0.25, 0.75], [1, 0.5], [0, 1]])
    assert np.array_equal(m["texcoords"][:3], want * np.float32([1, -1]))
    assert np.array_equal(m["texcoords"][3:], want)


# ---- alpha mask ---------------------------------------------------------------------------------------------
def half_alpha_texture(n=64):
    """alpha 0 for u < 0.5, 255 from there; colour bright"""
    t = np.full((n, n, 4), 230, np.uint8)
    t[:, : n // 2, 3] = 0
    t[:, n // 2:, 3] = 255
    return t


def masked_quad_scene(y, z0, z1, mode=vk.ALPHA_MASK, with_quad=True):
    """the rough plane of the closed-form NEE tests plus a one-sided quad at height y facing up, x in [-1, 1],
    z in [z0, z1], u = (x + 1) / 2, textured with half_alpha_texture"""
    from test_nee_oracle import plane_scene
    sc = plane_scene()
    if not with_quad:
        sc["texcoords"] = np.zeros((4, 2), np.float32)
        return sc
    q = np.array([[-1, y, z0], [1, y, z0], [1, y, z1], [-1, y, z1]], np.float32)
    sc["positions"] = np.concatenate([sc["positions"], q])
    sc["normals"] = np.tile(np.float32([0, 1, 0]), (8, 1))
    sc["indices"] = np.concatenate([sc["indices"], np.array([4, 6, 5, 4, 7, 6], np.uint32)])
    sc["tri_object"] = np.array([0, 0, 1, 1], np.uint32)
    sc["object_material"] = np.array([0, 1], np.uint32)
    quad = vk.material(color=(0.3, 0.6, 0.9), color_texture=1, alpha_mode=mode, alpha_cutoff=0.5)
    sc["materials"] = np.concatenate([sc["materials"], np.array([quad], vk.MATERIAL_DTYPE)])
    uvq = np.stack([(q[:, 0] + 1) / 2, (q[:, 2] - z0) / (z1 - z0)], -1)
    sc["texcoords"] = np.concatenate([np.zeros((4, 2), np.float32), uvq]).astype(np.float32)
    sc["textures"] = [half_alpha_texture()]
    return sc


def overhead_camera():
    """straight down from y = 5 with a 16 degree field: the plane within 0.7 units of the origin"""
    return vk.make_camera(eye=(0.0, 5.0, 0.0), forward=(0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0),
                          fov_y=float(np.radians(16.0)))


SIZE = 24
SIDE = (slice(None), slice(0, SIZE // 2 - 2)), (slice(None), slice(SIZE // 2 + 2, SIZE))


def alpha_renders(render, scene_args, light, disable):
    out = {}
    for key, kw in (("masked", {}), ("opaque", {"mode": vk.ALPHA_OPAQUE}), ("none", {"with_quad": False})):
        out[key] = render(masked_quad_scene(*scene_args, **kw), [light], disable)
    return out


def oracle_render(sc, lights, disable):
    return orc.pt_render(sc, overhead_camera(), SIZE, SIZE, 4, 2, 7, lights=lights, disable_material=disable)[..., :3]


def check_camera_ray_cutout(r, disable):
    """the quad over the whole view, light straight down: where the quad's alpha is cut out the image is the
    one without the quad, bit for bit; elsewhere the opaque quad's. With disable_material the camera rays keep
    the whole quad (stand-in alpha 1)"""
    left, right = SIDE
    assert np.array_equal(r["masked"][right], r["opaque"][right])
    if disable:
        assert np.array_equal(r["masked"][left], r["opaque"][left])
    else:
        assert np.array_equal(r["masked"][left], r["none"][left])
        assert not np.array_equal(r["none"][left], r["opaque"][left])


def check_shadow_ray_cutout(r):
    """the quad outside the view, light from 45 degrees behind it: plane points whose shadow ray crosses the
    cut-out half are lit as without the quad, the others shadowed as by the opaque quad, whatever
    disable_material says"""
    left, right = SIDE
    assert np.array_equal(r["masked"][left], r["none"][left])
    assert np.array_equal(r["masked"][right], r["opaque"][right])
    assert r["opaque"][right].max() < 0.1 * r["none"][right].min()


# at y = 3 the camera sees |z| < 0.28; shadow rays toward SLANT cross y = 3 at z in [2.3, 3.7]
SHADOW_QUAD = (3.0, 2.0, 4.0)
DOWN = vk.light(vk.LIGHT_DIRECTIONAL, direction=(0, -1, 0), intensity=3.0)
SLANT = vk.light(vk.LIGHT_DIRECTIONAL, direction=(0, -np.sqrt(0.5), -np.sqrt(0.5)), intensity=3.0)


@pytest.mark.parametrize("disable", [0, 1])
def test_alpha_mask_camera_rays(disable):
    check_camera_ray_cutout(alpha_renders(oracle_render, (1.0, -1.0, 1.0), DOWN, disable), disable)


@pytest.mark.parametrize("disable", [0, 1])
def test_alpha_mask_shadow_rays(disable):
    check_shadow_ray_cutout(alpha_renders(oracle_render, SHADOW_QUAD, SLANT, disable))


# ---- ray cone level of detail -------------------------------------------------------------------------------
LOD_W, LOD_H, LOD_FOV, LOD_SEED = 40, 32, 20.0, 3
TEX_N = 256


def lod_texture():
    rng = np.random.default_rng(11)
    t = rng.integers(96, 256, (TEX_N, TEX_N, 4), dtype=np.uint8)
    t[..., 3] = 255
    return t


def lod_scene(textured):
    """a thin index-matched slab (z = -2 and -2.1, x < 0) in front of a textured mirror at z = -5 whose
    coordinates span 4 units per unit; camera at the origin looking down -z. Rays through the slab reach the
    mirror with a cone widened over two segments (and bent by the integrator's eta bias of 1e-4 at each
    face); the others reach it first"""
    pos, idx, tri_obj, uv = [], [], [], []

    def quad(corners, obj, tex=None):
        base = len(pos)
        pos.extend(corners)
        uv.extend(tex if tex is not None else [[0, 0]] * 4)
        idx.extend([base, base + 1, base + 2, base, base + 2, base + 3])
        tri_obj.extend([obj, obj])

    quad([[-10, -10, -2], [0, -10, -2], [0, 10, -2], [-10, 10, -2]], 0)          # slab front, facing +z
    quad([[-10, -10, -2.1], [-10, 10, -2.1], [0, 10, -2.1], [0, -10, -2.1]], 0)  # slab back, facing -z
    corners = [[-2, -2, -5], [2, -2, -5], [2, 2, -5], [-2, 2, -5]]
    quad(corners, 1, [[(c[0] + 2) / 4, (c[1] + 2) / 4] for c in corners])
    pos = np.array(pos, np.float32)
    tris = np.array(idx).reshape(-1, 3)
    p = pos[tris]
    fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    nrm = np.zeros_like(pos)
    for t, tri in enumerate(tris):
        nrm[tri] = fn[t] / np.linalg.norm(fn[t])
    slab = vk.material(color=(1, 1, 1), roughness=0.0, transmission=1.0, ior=1.0)
    mirror = vk.material(color=(1, 1, 1), roughness=0.0, metalness=1.0, color_texture=1 if textured else 0)
    return {"positions": pos, "normals": nrm.astype(np.float32), "indices": np.array(idx, np.uint32),
            "tri_object": np.array(tri_obj, np.uint32), "object_material": np.array([0, 1], np.uint32),
            "materials": np.array([slab, mirror], vk.MATERIAL_DTYPE), "environment": 1.0,
            "texcoords": np.array(uv, np.float32), "textures": [lod_texture()]}


def lod_camera():
    return vk.make_camera(eye=(0.0, 0.0, 0.0), forward=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0),
                          fov_y=float(np.radians(LOD_FOV)))


def lod_lookups(seed=LOD_SEED):
    """per pixel: where the mirror is hit (u, v), the level of detail and whether the ray came through the
    slab. The camera rays are rebuilt from the pixel seeds (one sample: jitter = frac(2 rnd - 1) on both
    axes); the cone's spread and widths follow from the geometry, rounded to halves after each segment as
    the payload keeps them; lambda = 0.5 log2(uv area / world area) + log2(width) + 0.5 log2(n n) - log2(N.V)
    with the width from before the mirror's segment (0 for a first hit)"""
    half = lambda x: np.float64(np.float16(x))
    _, _, _, _, thx, thy = W.camera_basis(lod_camera(), LOD_W, LOD_H)
    spread = half(np.arctan(2.0 * np.tan(np.radians(LOD_FOV) / 2.0) / LOD_H))
    ys, xs = np.mgrid[0:LOD_H, 0:LOD_W]
    jit = np.zeros((LOD_H, LOD_W))
    for y in range(LOD_H):
        for x in range(LOD_W):
            off = 2.0 * float(orc.rnd_sequence(orc.xxhash32(seed, LOD_W * y + x), 1)[0][0]) - 1.0
            jit[y, x] = off - np.floor(off)
    dx = 2.0 * (xs + jit) / LOD_W - 1.0
    dy = 2.0 * (ys + jit) / LOD_H - 1.0
    d = np.stack([dx * float(thx), -dy * float(thy), -np.ones_like(dx)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    through = d[..., 0] < 0  # crosses z = -2 at x < 0

    def refract(i, eta=1.0 + 1e-4):
        # both faces see eta = 1 / 1 + 1e-4 (the integrator's eta bias), normal +z facing the ray
        n = np.array([0.0, 0.0, 1.0])
        ndi = i @ n
        k = 1.0 - eta * eta * (1.0 - ndi * ndi)
        r = eta * i - (eta * ndi + np.sqrt(k))[..., None] * n
        return r / np.linalg.norm(r, axis=-1, keepdims=True)

    d1 = refract(d)
    d2 = refract(d1)
    t1, t2 = 2.0 / -d[..., 2], 0.1 / -d1[..., 2]
    t3 = 2.9 / -d2[..., 2]
    hit_through = d * t1[..., None] + d1 * t2[..., None] + d2 * t3[..., None]
    hit_direct = d * (5.0 / -d[..., 2])[..., None]
    hit = np.where(through[..., None], hit_through, hit_direct)
    u, v = (hit[..., 0] + 2) / 4, (hit[..., 1] + 2) / 4
    w1 = np.vectorize(half)(spread * t1)
    w2 = np.vectorize(half)(w1 + spread * t2)
    nov = np.where(through, -d2[..., 2], -d[..., 2])
    with np.errstate(divide="ignore"):
        lam = np.where(through, 0.5 * np.log2(1 / 16) + np.log2(w2) + np.log2(TEX_N) - np.log2(nov), -np.inf)
    return u, v, lam, through


def lod_expected(level0=False, seed=LOD_SEED):
    u, v, lam, through = lod_lookups(seed)
    lam = np.zeros_like(lam) if level0 else lam
    val = sample(mip_chain(lod_texture()), u.ravel(), v.ravel(), lam.ravel())
    return val[:, :3].reshape(LOD_H, LOD_W, 3), through, lam


def oracle_lod(sc):
    return orc.pt_render(sc, lod_camera(), LOD_W, LOD_H, 1, 7, LOD_SEED)[..., :3]


def check_lod(render, seed=LOD_SEED):
    """the mirror reflects the sky with its base colour (metal, normal incidence): the image with the texture
    over the one without is the texture value at the pixel's lookup"""
    ratio = render(lod_scene(True)) / render(lod_scene(False))
    want, through, lam = lod_expected(seed=seed)
    # the case covers lookups between levels 0 and 2 through the slab and level 0 for first hits
    assert lam[through].min() > 0.2 and lam[through].max() < 2.5 and (~through).sum() > 100
    err = np.abs(ratio - want).max(axis=2)
    # (a width on a half's rounding boundary can land one half-ulp off: 7e-4 in lambda)
    assert err[~through].max() < 1e-4 and err[through].max() < 5e-4, (err[through].max(), err[~through].max())
    assert np.percentile(err[through], 95) < 1e-4
    # a lost cone width (level 0 through the slab) would miss by far more
    flat, _, _ = lod_expected(level0=True, seed=seed)
    assert np.abs(ratio - flat).max(axis=2)[through].mean() > 0.03


def test_ray_cone_level_of_detail():
    """the mirror's texture seen through the slab is sampled at the cone's level of detail; seen directly
    (cone width 0) at level 0"""
    check_lod(oracle_lod)


def test_gltf_base_color_texture_alpha_mode_and_texture_transform(tmp_path):
    """a glTF material with an embedded PNG base color texture, MASK mode with its cutoff, and
    KHR_texture_transform (offset, rotation about -z, scale: gltf.cpp:162-194) folded into the coordinates:
    u' = cos sx u + sin sy v + ox, v' = -sin sx u + cos sy v + oy"""
    import base64
    import json
    from PIL import Image
    import io
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (5, 7, 4), dtype=np.uint8)
    buf = io.BytesIO()
    Image.fromarray(img, "RGBA").save(buf, format="PNG")
    png = buf.getvalue()
    pos = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    uv = np.float32([[0.1, 0.2], [0.9, 0.3], [0.4, 0.8]])
    data = pos.tobytes() + uv.tobytes() + np.uint16([0, 1, 2, 0]).tobytes() + png
    data += b"\0" * (-len(data) % 4)
    off, rot, scale = [0.25, -0.5], 0.7, [2.0, 3.0]
    doc = {"asset": {"version": "2.0"}, "scenes": [{"nodes": [0, 1]}],
           "nodes": [{"mesh": 0}, {"mesh": 1}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": 2, "material": 0}]},
                      {"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": 2, "material": 1}]}],
           "materials": [{"alphaMode": "MASK", "alphaCutoff": 0.3,
                          "pbrMetallicRoughness": {"baseColorTexture": {
                              "index": 0, "extensions": {"KHR_texture_transform": {
                                  "offset": off, "rotation": rot, "scale": scale}}}}},
                         {"alphaMode": "BLEND", "pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}],
           "textures": [{"source": 0}], "images": [{"bufferView": 3, "mimeType": "image/png"}],
           "buffers": [{"byteLength": len(data),
                        "uri": "data:application/octet-stream;base64," + base64.b64encode(data).decode()}],
           "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36},
                           {"buffer": 0, "byteOffset": 36, "byteLength": 24},
                           {"buffer": 0, "byteOffset": 60, "byteLength": 6},
                           {"buffer": 0, "byteOffset": 68, "byteLength": len(png)}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                         {"bufferView": 1, "componentType": 5126, "count": 3, "type": "VEC2"},
                         {"bufferView": 2, "componentType": 5123, "count": 3, "type": "SCALAR"}]}
    (tmp_path / "t.gltf").write_text(json.dumps(doc))
    m = vk.load_model(str(tmp_path / "t.gltf"))
    assert len(m["textures"]) == 1 and np.array_equal(m["textures"][0], img)
    a, b = m["materials"][:2]
    assert a["color_texture"] == 1 and a["alpha_mode"] == vk.ALPHA_MASK and a["alpha_cutoff"] == np.float32(0.3)
    assert b["color_texture"] == 1 and b["alpha_mode"] == vk.ALPHA_BLEND
    c, s = np.cos(rot), np.sin(rot)
    u, v = uv[:, 0].astype(np.float64), uv[:, 1].astype(np.float64)
    want = np.stack([c * scale[0] * u + s * scale[1] * v + off[0], -s * scale[0] * u + c * scale[1] * v + off[1]], -1)
    assert np.allclose(m["texcoords"][:3], want, atol=1e-6)
    assert np.array_equal(m["texcoords"][3:], uv)  # no transform on the second material


# ---- alpha blend --------------------------------------------------------------------------------------------
def blend_quads_scene(y_levels, z0, z1, alpha, mode=vk.ALPHA_BLEND):
    """the rough plane plus one-sided quads facing up at the given heights (x in [-1, 1], z in [z0, z1]),
    untextured, color alpha `alpha`"""
    from test_nee_oracle import plane_scene
    sc = plane_scene()
    pos, idx, tobj = [sc["positions"]], [sc["indices"]], [0, 0]
    for k, y in enumerate(y_levels):
        base = 4 * (k + 1)
        pos.append(np.array([[-1, y, z0], [1, y, z0], [1, y, z1], [-1, y, z1]], np.float32))
        idx.append(np.array([base, base + 2, base + 1, base, base + 3, base + 2], np.uint32))
        tobj += [1, 1]
    sc["positions"] = np.concatenate(pos)
    sc["normals"] = np.tile(np.float32([0, 1, 0]), (len(sc["positions"]), 1))
    sc["indices"] = np.concatenate(idx)
    sc["tri_object"] = np.array(tobj, np.uint32)
    sc["object_material"] = np.array([0, 1], np.uint32)
    quad = vk.material(color=(0.3, 0.6, 0.9), alpha=alpha, alpha_mode=mode)
    sc["materials"] = np.concatenate([sc["materials"], np.array([quad], vk.MATERIAL_DTYPE)])
    return sc


def blend_threshold(seed, x, y, width):
    """rnd() of a copy of the payload state at the first trace of a pixel's first sample: the pixel state
    xxhash32(seed, W y + x) after the sample offset and the two lens draws, hashed with depth 0"""
    _, s = orc.rnd_sequence(orc.xxhash32(seed, width * y + x), 3)
    return orc.rnd_sequence(orc.xxhash32(s, 0), 1)[0][0]


def check_blend_camera_rays(render, seed):
    """one sample per pixel, the BLEND quad (alpha 0.4) over the whole view under the sky, no lights: a pixel
    shows the quad exactly when its trace's draw is at most 0.4, and then equals the opaque render, else the
    render without the quad (bounce rays meet only the quad's culled back)"""
    a = np.float32(0.4)

    def scene(mode):
        sc = blend_quads_scene([1.0], -1.0, 1.0, a, mode)
        sc["environment"] = 1.0
        return sc
    r = {k: render(scene(mode), m) for k, (mode, m) in
         {"blend": (vk.ALPHA_BLEND, 1), "opaque": (vk.ALPHA_OPAQUE, 1), "none": (vk.ALPHA_BLEND, 0)}.items()}
    keep = np.array([[not (a < blend_threshold(seed, x, y, SIZE)) for x in range(SIZE)] for y in range(SIZE)])
    assert 0.2 < keep.mean() < 0.6
    assert np.array_equal(r["blend"][keep], r["opaque"][keep])
    assert np.array_equal(r["blend"][~keep], r["none"][~keep])


def test_alpha_blend_camera_rays():
    def render(sc, with_quad):
        if not with_quad:
            sc["materials"][1]["color"][3] = 0.0  # alpha 0 < any draw: the quad is never kept
        return orc.pt_render(sc, overhead_camera(), SIZE, SIZE, 1, 2, 5)[..., :3]
    check_blend_camera_rays(render, 5)


def check_blend_shadow_rays(render):
    """BLEND occluders between the plane and a light straight down, out of the camera's view: each layer
    passes a shadow ray with probability 1 - alpha, independently (the per-candidate draws), so the lit image
    is (1 - alpha)^layers of the unoccluded one"""
    a = 0.3
    free = render(blend_quads_scene([], -3.0, 3.0, a)).mean()
    one = render(blend_quads_scene([3.0], -3.0, 3.0, a)).mean()
    two = render(blend_quads_scene([3.0, 3.5], -3.0, 3.0, a)).mean()
    assert abs(one / free - (1 - a)) < 0.01, (one / free, 1 - a)
    assert abs(two / free - (1 - a) ** 2) < 0.01, (two / free, (1 - a) ** 2)


BLEND_CAM = vk.make_camera(eye=(0.0, 0.5, 0.0), forward=(0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0),
                           fov_y=float(np.radians(16.0)))


def test_alpha_blend_shadow_rays():
    # the camera half a unit above the plane sees none of the quads at y = 3 and 3.5; every shadow ray from
    # its view goes straight up through them
    def render(sc):
        return orc.pt_render(sc, BLEND_CAM, 16, 16, 256, 2, 3, lights=[DOWN])[..., :3]
    check_blend_shadow_rays(render)


# ---- the other texture slots --------------------------------------------------------------------------------
def const_texture(rgba, size=(4, 6)):
    return np.tile(np.array(rgba, np.uint8), (*size, 1))


def with_planar_uv(sc):
    p = sc["positions"]
    sc["texcoords"] = np.stack([0.37 * p[:, 0] + 0.11 * p[:, 1], 0.41 * p[:, 2] + 0.23 * p[:, 1]], -1).astype(np.float32)
    return sc


SLOT_CASES = {
    # slot: (texel, material edit equivalent to the texture at that texel)
    "ao_rough_metal_texture": ((0, 128, 64, 255),
                               lambda m, t: (m.__setitem__("roughness", m["roughness"] * np.float32(t[1] / np.float32(255))),
                                             m.__setitem__("metalness", m["metalness"] * np.float32(t[2] / np.float32(255))))),
    "transmission_texture": ((200, 0, 0, 255),
                             lambda m, t: m.__setitem__("transmission", m["transmission"] * np.float32(t[0] / np.float32(255)))),
    "diffuse_transmission_texture": ((0, 0, 0, 170),
                                     lambda m, t: m.__setitem__("diffuse_transmission", m["diffuse_transmission"] * np.float32(t[3] / np.float32(255)))),
    "diffuse_transmission_color_texture": ((51, 102, 204, 0),
                                           lambda m, t: m.__setitem__("diffuse_transmission_color", np.concatenate(
                                               [m["diffuse_transmission_color"][:3] * (np.float32(t[:3]) / np.float32(255)), [0]]))),
}


def slot_scenes(slot):
    """the room under its sky with a constant texture in one slot of every material, and the same room with
    each material's factor multiplied by the texel instead: the two render bit for bit alike"""
    texel, edit = SLOT_CASES[slot]
    a = with_planar_uv(W.room_scene())
    a["materials"] = a["materials"].copy()
    if slot.startswith("diffuse_transmission"):
        a["materials"]["diffuse_transmission"] = 0.5
        a["materials"]["diffuse_transmission_color"] = (0.9, 0.8, 0.7, 0.0)
    if slot == "transmission_texture":
        a["materials"]["transmission"] = np.where(a["materials"]["transmission"] > 0, a["materials"]["transmission"], 0.5)
    b = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in a.items()}
    a["textures"] = [const_texture(texel)]
    a["materials"][slot] = 1
    for m in b["materials"]:
        edit(m, texel)
    return a, b


@pytest.mark.parametrize("slot", list(SLOT_CASES))
def test_texture_slots_scale_their_factors(slot):
    a, b = slot_scenes(slot)
    ra = orc.pt_render(a, W.room_camera(), 64, 36, 2, 6, 3)
    rb = orc.pt_render(b, W.room_camera(), 64, 36, 2, 6, 3)
    assert np.array_equal(ra, rb)


def emissive_quad(emission, strength, texel):
    """a quad facing a camera straight above it, emitting; the emission texture constant"""
    from test_nee_oracle import plane_scene
    sc = plane_scene(half=2.0)
    sc["materials"] = np.array([vk.material(emission=emission, emission_strength=strength, emission_texture=0)],
                               vk.MATERIAL_DTYPE)
    sc["materials"][0]["emission_texture"] = 1
    sc["texcoords"] = np.float32([[0, 0], [1, 0], [1, 1], [0, 1]])
    sc["textures"] = [const_texture(texel)]
    return sc


def test_emission_texture_multiplies_or_replaces():
    """seen directly (one segment), the radiance is the emission: the texel times a coloured emission, or the
    texel alone where the material's emission is black (raypipeline.slang:556-566), times the strength"""
    from test_nee_oracle import down_camera
    t = (51, 102, 153, 255)
    tex = np.float32(t[:3]) / np.float32(255)
    img = orc.pt_render(emissive_quad((0.5, 1.0, 2.0), 3.0, t), down_camera(), 4, 4, 2, 1, 1)
    want = (np.float32([0.5, 1.0, 2.0]) * tex) * np.float32(3.0)
    assert np.array_equal(img[..., :3], np.broadcast_to(want, img[..., :3].shape))
    img = orc.pt_render(emissive_quad((0.0, 0.0, 0.0), 3.0, t), down_camera(), 4, 4, 2, 1, 1)
    assert np.array_equal(img[..., :3], np.broadcast_to(tex * np.float32(3.0), img[..., :3].shape))


def normal_map_scenes(texel):
    """a flat plane with a constant normal map, and the plane tilted so that its geometric normal is the
    mapped one. The tangent follows Geometry::compute_tangents' sign (for u along +x, v along +z and the
    normal +y: dot(cross(n, t), b) = dot(-z, +z) < 0 keeps t = +x) and the bitangent is cross(n, t) = -z.
    Both quads are shifted so that their diagonals (where the triangle test may leave a crack) lie far
    outside the view"""
    from test_nee_oracle import plane_scene
    flat = plane_scene()
    flat["positions"] = flat["positions"] + np.float32([20, 0, -20])
    flat["texcoords"] = (flat["positions"][:, [0, 2]] / 10.0).astype(np.float32)
    flat["tangents"] = np.tile(np.float32([1, 0, 0]), (4, 1))
    flat["textures"] = [const_texture(texel)]
    flat["materials"]["normal_texture"] = 1
    c = np.float32(texel[:2]) / np.float32(255)
    nt = np.array([2 * (c[0] - 0.5), 2 * (c[1] - 0.5), 0.0])
    nt[2] = np.sqrt(max(0.0, 1 - nt[0] ** 2 - nt[1] ** 2))
    T, B, N = np.array([1.0, 0, 0]), np.array([0, 0, -1.0]), np.array([0, 1.0, 0])
    n = nt[0] * T + nt[1] * B + nt[2] * N
    n /= np.linalg.norm(n)
    # a plane through the origin with normal n, spanned by two unit vectors orthogonal to it
    e1 = np.cross(n, [0, 0, 1.0]); e1 /= np.linalg.norm(e1)
    e2 = np.cross(e1, n)
    c = 20 * (e1 - e2)
    corners = np.array([c - 50 * e1 - 50 * e2, c + 50 * e1 - 50 * e2, c + 50 * e1 + 50 * e2, c - 50 * e1 + 50 * e2],
                       np.float32)
    tilted = plane_scene()
    tilted["positions"] = corners
    tilted["normals"] = np.tile(n.astype(np.float32), (4, 1))
    tilted["indices"] = np.array([0, 1, 2, 0, 2, 3], np.uint32) if np.dot(np.cross(corners[1] - corners[0], corners[2] - corners[0]), n) > 0 \
        else np.array([0, 2, 1, 0, 3, 2], np.uint32)
    return flat, tilted, n


@pytest.mark.parametrize("texel", [(160, 128, 255, 255), (90, 200, 255, 255), (128, 128, 255, 255)])
def test_normal_map_tilts_like_geometry(texel):
    """the normal-mapped plane shades as the plane tilted to the mapped normal: NEE from a light straight down
    is the same function of (shading normal, view, light) for both, to rounding"""
    from test_nee_oracle import down_camera
    flat, tilted, n = normal_map_scenes(texel)
    light = vk.light(vk.LIGHT_DIRECTIONAL, direction=(0, -1, 0), intensity=2.0)
    a = orc.pt_render(flat, down_camera(), 8, 8, 64, 2, 4, lights=[light])[..., :3]
    b = orc.pt_render(tilted, down_camera(), 8, 8, 64, 2, 4, lights=[light])[..., :3]
    assert np.allclose(a, b, rtol=1e-5, atol=1e-7), (np.abs(a - b).max(), n)
    flat["materials"]["normal_texture"] = 0
    c = orc.pt_render(flat, down_camera(), 8, 8, 64, 2, 4, lights=[light])[..., :3]
    assert np.abs(a - c).max() > 0.01 * a.mean()


def test_obj_tangents_follow_compute_tangents(tmp_path):
    """a quad in the xz plane, u along +x and v along +z, normal +y: the s direction is +x and the handedness
    test dot(cross(n, t), b) = dot(-z, +z) < 0 keeps its sign (Geometry.cpp:242-253); a normal map in MTL's
    `norm` fills the material's normal slot"""
    Image = pytest.importorskip("PIL.Image")
    Image.fromarray(np.full((2, 2, 4), 128, np.uint8)).save(tmp_path / "n.png")
    (tmp_path / "q.mtl").write_text("newmtl m\nnorm n.png\n")
    (tmp_path / "q.obj").write_text("mtllib q.mtl\nv 0 0 0\nv 1 0 0\nv 1 0 1\nv 0 0 1\nvt 0 0\nvt 1 0\nvt 1 1\n"
                                    "vt 0 1\nvn 0 1 0\nusemtl m\nf 1/1/1 4/4/1 3/3/1 2/2/1\n")
    m = vk.load_model(str(tmp_path / "q.obj"))
    assert np.array_equal(m["tangents"], np.tile(np.float32([1, 0, 0]), (6, 1)))
    assert m["materials"][0]["normal_texture"] == 1 and m["materials"][0]["color_texture"] == 0
    # no v flip without a diffuse texture
    assert np.array_equal(m["texcoords"][:3], np.float32([[0, 0], [0, 1], [1, 1]]))


def test_gltf_texture_slots_and_tangents(tmp_path):
    """glTF texture slots (metallicRoughness, normal, emissive with the factor zeroed, KHR transmission and
    diffuse transmission textures) and a TANGENT attribute (xyz kept, w dropped) turned by the node's
    rotation; a primitive without TANGENT gets compute_tangents"""
    import base64
    import io
    import json
    from PIL import Image
    buf = io.BytesIO()
    Image.fromarray(np.full((2, 2, 4), 200, np.uint8), "RGBA").save(buf, format="PNG")
    png = buf.getvalue()
    pos = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    uv = np.float32([[0, 0], [1, 0], [0, 1]])
    nrm = np.float32([[0, 0, 1]] * 3)
    tg = np.float32([[1, 0, 0, -1]] * 3)
    data = pos.tobytes() + uv.tobytes() + nrm.tobytes() + tg.tobytes() + png
    data += b"\0" * (-len(data) % 4)
    views = [(0, 36), (36, 24), (60, 36), (96, 48), (144, len(png))]
    tex = lambda i: {"index": i}
    doc = {"asset": {"version": "2.0"}, "scenes": [{"nodes": [0, 1]}],
           "nodes": [{"mesh": 0, "rotation": [0, 0.7071067811865476, 0, 0.7071067811865476]}, {"mesh": 1}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1, "NORMAL": 2, "TANGENT": 3},
                                       "material": 0}]},
                      {"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1, "NORMAL": 2}, "material": 0}]}],
           "materials": [{"pbrMetallicRoughness": {"metallicRoughnessTexture": tex(0)}, "normalTexture": tex(0),
                          "emissiveFactor": [1, 1, 1], "emissiveTexture": tex(0),
                          "extensions": {"KHR_materials_transmission": {"transmissionTexture": tex(0)},
                                         "KHR_materials_diffuse_transmission": {
                                             "diffuseTransmissionTexture": tex(0),
                                             "diffuseTransmissionColorTexture": tex(0)}}}],
           "textures": [{"source": 0}], "images": [{"bufferView": 4, "mimeType": "image/png"}],
           "buffers": [{"byteLength": len(data),
                        "uri": "data:application/octet-stream;base64," + base64.b64encode(data).decode()}],
           "bufferViews": [{"buffer": 0, "byteOffset": o, "byteLength": n} for o, n in views],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                         {"bufferView": 1, "componentType": 5126, "count": 3, "type": "VEC2"},
                         {"bufferView": 2, "componentType": 5126, "count": 3, "type": "VEC3"},
                         {"bufferView": 3, "componentType": 5126, "count": 3, "type": "VEC4"}]}
    (tmp_path / "s.gltf").write_text(json.dumps(doc))
    m = vk.load_model(str(tmp_path / "s.gltf"))
    mat = m["materials"][0]
    for slot in ("ao_rough_metal_texture", "normal_texture", "emission_texture", "transmission_texture",
                 "diffuse_transmission_texture", "diffuse_transmission_color_texture"):
        assert mat[slot] == 1, slot
    assert np.array_equal(mat["emission"][:3], [0, 0, 0])
    # rotation by 90 degrees about +y turns +x into -z
    assert np.allclose(m["tangents"][:3], np.tile([0, 0, -1], (3, 1)), atol=1e-7)
    # computed for the second primitive: s direction +x, normal +z, b = +y, cross(n, t) = +y -> sign -1
    assert np.array_equal(m["tangents"][3:], np.tile(np.float32([-1, 0, 0]), (3, 1)))


# ---- samplers -------------------------------------------------------------------------------------------------
SCIPY_MODE = {vk.ADDRESS_REPEAT: "grid-wrap", vk.ADDRESS_MIRRORED_REPEAT: "reflect", vk.ADDRESS_CLAMP_TO_EDGE: "nearest"}


def sample_level_modes(level, u, v, mode_u, mode_v, nearest):
    """one level under Vulkan's address modes via scipy (grid-wrap = repeat, reflect = mirrored repeat, nearest =
    clamp to edge; order 0 = nearest filtering at texel centres). Mixed modes per axis: sample a copy padded
    with the v-axis mode, then interpolate along u with its own."""
    h, w = level.shape[:2]
    x, y = u * w - 0.5, v * h - 0.5
    out = np.zeros((len(u), 4))
    for c in range(4):
        ch = level[..., c]
        if mode_u == mode_v:
            out[:, c] = ndimage.map_coordinates(ch, [y, x], order=0 if nearest else 1, mode=SCIPY_MODE[mode_u])
        else:
            # separable: linear (or nearest) along v with mode_v at the two (or one) columns around x
            if nearest:
                xi = np.floor(u * w).astype(np.int64)
                cols = np.array([_address(k, w, mode_u) for k in xi])
                colv = np.stack([ndimage.map_coordinates(ch[:, k], [yy], order=0, mode=SCIPY_MODE[mode_v])
                                 for k, yy in zip(cols, y[:, None])])[:, 0]
                out[:, c] = colv
            else:
                x0 = np.floor(x).astype(np.int64)
                ax = x - x0
                a = np.array([ndimage.map_coordinates(ch[:, _address(k, w, mode_u)], [[yy]], order=1,
                                                      mode=SCIPY_MODE[mode_v])[0] for k, yy in zip(x0, y)])
                b = np.array([ndimage.map_coordinates(ch[:, _address(k + 1, w, mode_u)], [[yy]], order=1,
                                                      mode=SCIPY_MODE[mode_v])[0] for k, yy in zip(x0, y)])
                out[:, c] = a + ax * (b - a)
    return out


def _address(i, n, mode):
    if mode == vk.ADDRESS_CLAMP_TO_EDGE:
        return min(max(i, 0), n - 1)
    if mode == vk.ADDRESS_MIRRORED_REPEAT:
        t = i % (2 * n)
        return t if t < n else 2 * n - 1 - t
    return i % n


@pytest.mark.parametrize("sampler", [(1, 1, 0, 0), (2, 2, 0, 0), (1, 2, 0, 0), (0, 0, 1, 1), (2, 1, 1, 1),
                                     (0, 1, 1, 0), (2, 0, 0, 1)])
def test_sampler_modes_match_scipy(sampler):
    """mirrored repeat (scipy's reflect: the edge texel repeats), clamp to edge (scipy's nearest extension) and
    repeat per axis; nearest filtering where the filter for the level of detail (min above 0, mag at or
    below) says so; the two levels around lambda blended linearly"""
    rng = np.random.default_rng(sum(sampler) + 7 * sampler[0])
    img = rng.integers(0, 256, (7, 12, 4), dtype=np.uint8)
    n = 600
    uv = rng.uniform(-2.5, 3.5, (n, 2)).astype(np.float32)
    lam = rng.uniform(-1.0, 3.0, n).astype(np.float32)
    got = orc.texture_sample(img, uv, lam, sampler)
    levels = mip_chain(img)
    u, v = uv[:, 0].astype(np.float64), uv[:, 1].astype(np.float64)
    lam_c = np.clip(lam.astype(np.float64), 0, len(levels) - 1)
    l0 = np.floor(lam_c).astype(int)
    f = lam_c - l0
    want = np.zeros((n, 4))
    for k in range(len(levels)):
        sel = l0 == k
        if not sel.any():
            continue
        near = np.where(lam[sel] > 0, sampler[2] == 1, sampler[3] == 1)
        for nearest in (False, True):
            s2 = sel.copy()
            s2[sel] = near == nearest
            if not s2.any():
                continue
            a = sample_level_modes(levels[k], u[s2], v[s2], sampler[0], sampler[1], nearest)
            b = sample_level_modes(levels[min(k + 1, len(levels) - 1)], u[s2], v[s2], sampler[0], sampler[1], nearest)
            want[s2] = a + f[s2][:, None] * (b - a)
    assert np.abs(got - want).max() < 2e-5


def test_gltf_samplers(tmp_path):
    """glTF sampler fields map as convert_sampler does (CLAMP_TO_EDGE 33071, MIRRORED_REPEAT 33648, NEAREST
    9728; anything else repeat / linear); one texture per (image, sampler) pair"""
    import base64
    import io
    import json
    from PIL import Image
    buf = io.BytesIO()
    Image.fromarray(np.full((2, 2, 4), 90, np.uint8), "RGBA").save(buf, format="PNG")
    png = buf.getvalue()
    pos = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    data = pos.tobytes() + png
    data += b"\0" * (-len(data) % 4)
    tex = lambda i: {"index": i}
    doc = {"asset": {"version": "2.0"}, "scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "material": 0}]}],
           "materials": [{"pbrMetallicRoughness": {"baseColorTexture": tex(0), "metallicRoughnessTexture": tex(1)},
                          "normalTexture": tex(2), "emissiveTexture": tex(0)}],
           "samplers": [{"wrapS": 33071, "wrapT": 33648, "minFilter": 9728, "magFilter": 9729},
                        {"minFilter": 9987, "magFilter": 9728}],
           "textures": [{"source": 0, "sampler": 0}, {"source": 0, "sampler": 1}, {"source": 0}],
           "images": [{"bufferView": 1, "mimeType": "image/png"}],
           "buffers": [{"byteLength": len(data),
                        "uri": "data:application/octet-stream;base64," + base64.b64encode(data).decode()}],
           "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36},
                           {"buffer": 0, "byteOffset": 36, "byteLength": len(png)}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"}]}
    (tmp_path / "s.gltf").write_text(json.dumps(doc))
    m = vk.load_model(str(tmp_path / "s.gltf"))
    mat = m["materials"][0]
    assert len(m["textures"]) == 3
    smp = m["texture_samplers"]
    assert smp[mat["color_texture"] - 1] == (vk.ADDRESS_CLAMP_TO_EDGE, vk.ADDRESS_MIRRORED_REPEAT, vk.FILTER_NEAREST,
                                            vk.FILTER_LINEAR)
    assert smp[mat["ao_rough_metal_texture"] - 1] == (0, 0, vk.FILTER_LINEAR, vk.FILTER_NEAREST)
    assert smp[mat["normal_texture"] - 1] == (0, 0, 0, 0)
    assert mat["emission_texture"] == mat["color_texture"]
