"""Synthetic scenes and ray batches of the benchmark configurations (host side, numpy).

All randomness comes from one documented generator: SplitMix64 used counter-style, draw k of a stream
with seed s being mix(s + (k + 1) * 0x9E3779B97F4A7C15); a uniform float is (draw >> 40) * 2^-24 and a
normal is Box-Muller on two consecutive uniforms, cos branch, evaluated in float64 and rounded to float32.
Meshes come from the library's Geometry::IcoSphere / Geometry::Plane (``vierkant_amd.icosphere`` /
``vierkant_amd.plane``), so triangle order is the reference's.

Rays are stored as vierkant_amd.RAY_DTYPE (origin, tmin, direction, tmax); directions are normalized on
the host with glm's order, d * (1 / sqrt((x*x + y*y) + z*z)), in float32, as the reference's Ray
constructor does (include/vierkant/intersection.hpp:152).
"""
import numpy as np

import vierkant_amd as vk

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1, M2 = np.uint64(0xBF58476D1CE4E5B9), np.uint64(0x94D049BB133111EB)


def sm64(seed, start, count):
    """draws start .. start+count-1 of the SplitMix64 stream `seed` (uint64 array)"""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, start, count):
    return (sm64(seed, start, count) >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)


def normal(seed, start, count):
    """count normals from uniforms start .. start + 2 count - 1 (pairs)"""
    u = uniform(seed, start, 2 * count).astype(np.float64).reshape(count, 2)
    r = np.sqrt(-2.0 * np.log(1.0 - u[:, 0]))
    return (r * np.cos(2.0 * np.pi * u[:, 1])).astype(np.float32)


def glm_normalize(d):
    """float32 rows d * (1 / sqrt((x*x + y*y) + z*z))"""
    d = np.asarray(d, dtype=np.float32)
    sq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    inv = np.float32(1.0) / np.sqrt(sq)
    return d * inv[:, None]


def make_rays(o, d, tmin=0.0, tmax=np.inf):
    n = len(o)
    r = np.zeros(n, dtype=vk.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = o[:, 0], o[:, 1], o[:, 2]
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def flat_triangles(pos, idx):
    """(n_tri, 9) float32 v0 v1 v2 rows of an indexed mesh"""
    return np.ascontiguousarray(pos[idx.reshape(-1, 3)].reshape(-1, 9).astype(np.float32))


def random_sphere_dirs(seed, start, count):
    """uniform directions on S^2 from 2 uniforms each: z = 1 - 2 u0, phi = 2 pi u1"""
    u = uniform(seed, start, 2 * count).reshape(count, 2).astype(np.float64)
    z = 1.0 - 2.0 * u[:, 0]
    phi = 2.0 * np.pi * u[:, 1]
    rxy = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], axis=1).astype(np.float32)
    return glm_normalize(d)


# ------------------------------------------------------------------------------------------------
# configuration 1: jittered icosphere + random rays (brute-force case)
# ------------------------------------------------------------------------------------------------
def jittered_icosphere(subdiv=5, radius=1.0, sigma=0.01, seed=0):
    """Geometry::IcoSphere(radius, subdiv), every vertex copy moved by N(0, sigma) per axis. The
    reference's tessellation keeps separate copies of shared edge vertices, so jitter opens cracks.
    Uses normals 0 .. 3 n_vert - 1 of the stream (uniform draws 0 .. 6 n_vert - 1)."""
    pos, idx = vk.icosphere(radius, subdiv)
    n = normal(seed, 0, pos.size).reshape(pos.shape)
    jit = (np.float64(sigma) * n.astype(np.float64)).astype(np.float32)
    return (pos + jit).astype(np.float32), idx


def icosphere_rays(n_rays=65536, subdiv=5, seed=0):
    """65,536 rays (origins U[-3,3]^3, directions uniform on S^2) vs the jittered subdiv-5 icosphere.
    Ray r uses uniform draws K0 + 5 r .. K0 + 5 r + 4 with K0 = 6 n_vert (after the jitter)."""
    pos, idx = jittered_icosphere(subdiv, seed=seed)
    k0 = 2 * pos.size
    u = uniform(seed, k0, 5 * n_rays).reshape(n_rays, 5)
    o = (np.float32(-3.0) + np.float32(6.0) * u[:, :3]).astype(np.float32)
    z = 1.0 - 2.0 * u[:, 3].astype(np.float64)
    phi = 2.0 * np.pi * u[:, 4].astype(np.float64)
    rxy = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = glm_normalize(np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], axis=1).astype(np.float32))
    return pos, idx, make_rays(o, d)


# ------------------------------------------------------------------------------------------------
# configuration 2: fBm heightfield + 1,000 icospheres, pinhole + bounce + shadow rays
# ------------------------------------------------------------------------------------------------
def _hash2(ix, iz, octave, seed):
    """32-bit lattice hash (murmur3 finalizer over a linear mix), value in [-1, 1]"""
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint32) * np.uint32(0x8da6b343)) ^ (iz.astype(np.uint32) * np.uint32(0xd8163841))
        h ^= np.uint32((octave * 0xcb1ab31f + seed * 0x165667b1) & 0xffffffff)
        h ^= h >> np.uint32(16)
        h *= np.uint32(0x85ebca6b)
        h ^= h >> np.uint32(13)
        h *= np.uint32(0xc2b2ae35)
        h ^= h >> np.uint32(16)
    return h.astype(np.float64) / 4294967295.0 * 2.0 - 1.0


def fbm(x, z, octaves=6, base_freq=0.4, lacunarity=2.0, gain=0.5, seed=0):
    """value-noise fBm (smoothstep-interpolated lattice values), float64"""
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    total = np.zeros_like(x)
    amp, freq = 1.0, base_freq
    for o in range(octaves):
        fx, fz = x * freq, z * freq
        ix, iz = np.floor(fx), np.floor(fz)
        tx, tz = fx - ix, fz - iz
        sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
        ix, iz = ix.astype(np.int64), iz.astype(np.int64)
        v00, v10 = _hash2(ix, iz, o, seed), _hash2(ix + 1, iz, o, seed)
        v01, v11 = _hash2(ix, iz + 1, o, seed), _hash2(ix + 1, iz + 1, o, seed)
        a = v00 + (v10 - v00) * sx
        b = v01 + (v11 - v01) * sx
        total += amp * (a + (b - a) * sz)
        amp *= gain
        freq *= lacunarity
    return total


def heightfield_spheres_scene(grid=512, n_spheres=1000, seed=0, amplitude=0.2, size=10.0):
    """Heightfield Plane(10, 10, 512, 512) displaced by 0.2 fBm (524,288 tris) followed by 1,000
    IcoSphere(r, 3) (1,280 tris each) with r = 0.01 * 10^U, centers x, z ~ U[-5, 5] and y = surface
    height + U[0, 1]; each sphere takes 4 uniform draws (r, x, z, y) from stream `seed`.
    Returns (positions [n, 3] float32, indices uint32)."""
    pos, idx = vk.plane(size, size, grid, grid)
    pos[:, 1] = (amplitude * fbm(pos[:, 0], pos[:, 2], seed=seed)).astype(np.float32)
    u = uniform(seed, 0, 4 * n_spheres).reshape(n_spheres, 4).astype(np.float64)
    radii = (0.01 * 10.0 ** u[:, 0]).astype(np.float32)
    cx = (-0.5 * size + size * u[:, 1])
    cz = (-0.5 * size + size * u[:, 2])
    cy = amplitude * fbm(cx, cz, seed=seed) + u[:, 3]
    centers = np.stack([cx, cy, cz], axis=1).astype(np.float32)
    all_pos, all_idx = [pos], [idx]
    base = len(pos)
    for s in range(n_spheres):
        sp, si = vk.icosphere(float(radii[s]), 3)
        all_pos.append((sp + centers[s]).astype(np.float32))
        all_idx.append((si + np.uint32(base)).astype(np.uint32))
        base += len(sp)
    return np.ascontiguousarray(np.concatenate(all_pos)), np.ascontiguousarray(np.concatenate(all_idx))


def heightfield_spheres_instanced(grid=512, n_spheres=1000, seed=0, amplitude=0.2, size=10.0):
    """the heightfield + spheres scene as a two-level scene: the displaced plane as one mesh, the spheres as
    instances of one IcoSphere(1, 3) scaled by r and translated to their centers (the same draws as
    heightfield_spheres_scene). Returns (meshes [(pos, idx), (pos, idx)], object-to-world 3x4 matrices,
    mesh index per instance)."""
    pos, idx = vk.plane(size, size, grid, grid)
    pos[:, 1] = (amplitude * fbm(pos[:, 0], pos[:, 2], seed=seed)).astype(np.float32)
    u = uniform(seed, 0, 4 * n_spheres).reshape(n_spheres, 4).astype(np.float64)
    radii = (0.01 * 10.0 ** u[:, 0]).astype(np.float32)
    cx = (-0.5 * size + size * u[:, 1])
    cz = (-0.5 * size + size * u[:, 2])
    cy = amplitude * fbm(cx, cz, seed=seed) + u[:, 3]
    centers = np.stack([cx, cy, cz], axis=1).astype(np.float32)
    unit = vk.icosphere(1.0, 3)
    mats = [np.hstack([np.eye(3), np.zeros((3, 1))])]
    mesh = [0]
    for s in range(n_spheres):
        m = np.zeros((3, 4), np.float64)
        m[:, :3] = np.eye(3) * float(radii[s])
        m[:, 3] = centers[s]
        mats.append(m)
        mesh.append(1)
    return [(pos, idx), unit], mats, mesh


def pinhole_rays(width=2048, height=2048, eye=(0.0, 3.0, 7.0), target=(0.0, 0.0, 0.0), fov_y_deg=50.0):
    """one ray through each pixel center, row-major from the top-left, tmin 0, tmax inf"""
    eye = np.asarray(eye, dtype=np.float64)
    fwd = np.asarray(target, dtype=np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = np.tan(np.radians(fov_y_deg) * 0.5)
    aspect = width / height
    xs = ((np.arange(width) + 0.5) / width * 2.0 - 1.0) * th * aspect
    ys = (1.0 - (np.arange(height) + 0.5) / height * 2.0) * th
    px, py = np.meshgrid(xs, ys)
    d = fwd[None, :] + px.reshape(-1, 1) * right[None, :] + py.reshape(-1, 1) * up[None, :]
    d = glm_normalize(d.astype(np.float32))
    o = np.broadcast_to(eye.astype(np.float32), d.shape)
    return make_rays(np.ascontiguousarray(o), d)


def secondary_rays(primary, hits, tris_flat, seed=0, light=(0.3, 1.0, 0.2), eps=1e-4):
    """Cosine-hemisphere bounce rays and directional shadow rays from the primary hits.
    Hit point p = o + t d, geometric normal n = normalize(cross(e1, e2)) facing the incoming ray, origin
    p + eps n (tmin = eps). Bounce ray i uses uniform draws 2 i, 2 i + 1 of stream seed + 1. A primary
    miss keeps its own ray as the bounce ray (it misses again) and sends its shadow ray from the
    primary origin toward the light."""
    n = len(primary)
    hit = hits["prim"] != vk.INVALID_PRIM
    o = np.stack([primary["ox"], primary["oy"], primary["oz"]], axis=1)
    d = np.stack([primary["dx"], primary["dy"], primary["dz"]], axis=1)
    t = np.where(hit, hits["t"], 0.0).astype(np.float32)
    p = (o + t[:, None] * d).astype(np.float32)
    prim = np.where(hit, hits["prim"], 0).astype(np.int64)
    tri = tris_flat[prim].astype(np.float64)
    e1 = tri[:, 3:6] - tri[:, 0:3]
    e2 = tri[:, 6:9] - tri[:, 0:3]
    nrm = np.cross(e1, e2)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    flip = np.sum(nrm * d, axis=1) > 0
    nrm[flip] *= -1.0
    origin = (p + eps * nrm).astype(np.float32)
    # bounce: cosine-weighted hemisphere around n
    u = uniform(seed + 1, 0, 2 * n).reshape(n, 2).astype(np.float64)
    r, phi = np.sqrt(u[:, 0]), 2.0 * np.pi * u[:, 1]
    lx, ly, lz = r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - u[:, 0]))
    helper = np.where(np.abs(nrm[:, 0:1]) > 0.9, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    tx = np.cross(helper, nrm)
    tx /= np.linalg.norm(tx, axis=1, keepdims=True)
    ty = np.cross(nrm, tx)
    bd = glm_normalize((lx[:, None] * tx + ly[:, None] * ty + lz[:, None] * nrm).astype(np.float32))
    b_o = np.where(hit[:, None], origin, o).astype(np.float32)
    b_d = np.where(hit[:, None], bd, d).astype(np.float32)
    bounce = make_rays(b_o, b_d, tmin=0.0)
    bounce["tmin"] = np.where(hit, np.float32(eps), np.float32(0.0))
    # shadow: toward a directional light
    L = glm_normalize(np.asarray([light], dtype=np.float32))[0]
    s_o = np.where(hit[:, None], origin, o).astype(np.float32)
    shadow = make_rays(s_o, np.broadcast_to(L, (n, 3)).astype(np.float32), tmin=0.0)
    shadow["tmin"] = np.where(hit, np.float32(eps), np.float32(0.0))
    return bounce, shadow


# ------------------------------------------------------------------------------------------------
# configuration 3: Cornell-style room + 200 spheres + the jittered icosphere (path tracer)
# ------------------------------------------------------------------------------------------------
class _SceneBuilder:
    def __init__(self):
        self.pos, self.nrm, self.idx, self.obj, self.obj_mat, self.mats = [], [], [], [], [], []
        self.n_vert = 0

    def material(self, m):
        self.mats.append(m)
        return len(self.mats) - 1

    def add(self, pos, nrm, idx, mat):
        o = len(self.obj_mat)
        self.obj_mat.append(mat)
        self.pos.append(np.asarray(pos, np.float32))
        self.nrm.append(np.asarray(nrm, np.float32))
        self.idx.append(np.asarray(idx, np.uint32) + np.uint32(self.n_vert))
        self.obj.append(np.full(len(idx) // 3, o, np.uint32))
        self.n_vert += len(pos)
        return o

    def quads(self, quads, mat):
        """quads: list of (4 corners, outward normal); each becomes 2 counter-clockwise triangles facing n"""
        pos, nrm, idx = [], [], []
        for corners, n in quads:
            c = np.asarray(corners, np.float64)
            n = np.asarray(n, np.float64)
            if np.dot(np.cross(c[1] - c[0], c[2] - c[0]), n) < 0:
                c = c[::-1]
            b = len(pos)
            pos.extend(c)
            nrm.extend([n] * 4)
            idx.extend([b, b + 1, b + 2, b, b + 2, b + 3])
        return self.add(pos, nrm, idx, mat)

    def scene(self, environment=1.0):
        import vierkant_amd as vk
        return {"positions": np.ascontiguousarray(np.concatenate(self.pos)),
                "normals": np.ascontiguousarray(np.concatenate(self.nrm)),
                "indices": np.ascontiguousarray(np.concatenate(self.idx)),
                "tri_object": np.ascontiguousarray(np.concatenate(self.obj)),
                "object_material": np.asarray(self.obj_mat, np.uint32),
                "materials": np.array(self.mats, dtype=vk.MATERIAL_DTYPE), "environment": environment}


def _box_quads(lo, hi, inward):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    s = -1.0 if inward else 1.0
    return [([(x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)], (0, -s, 0)),
            ([(x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1)], (0, s, 0)),
            ([(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)], (0, 0, -s)),
            ([(x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)], (0, 0, s)),
            ([(x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1)], (-s, 0, 0)),
            ([(x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)], (s, 0, 0))]


def _random_material(u):
    """roughness U(0.05, 1), metallic Bernoulli(0.3), 10% transmissive (transmission 1, IOR 1.5); u = 6 uniforms"""
    import vierkant_amd as vk
    color = tuple(float(0.2 + 0.7 * x) for x in u[:3])
    rough = float(0.05 + 0.95 * u[3])
    metal = 1.0 if u[4] < 0.3 else 0.0
    trans = 1.0 if u[5] < 0.1 else 0.0
    return vk.material(color=color, roughness=rough, metalness=metal, transmission=trans, ior=1.5)


def room_scene(n_spheres=200, seed=0):
    """Room x in [-3, 3], y in [0, 6], z in [-3, 3] with the front (z = +3) open to the sky: floor, ceiling,
    back, red left and green right walls, an emissive ceiling quad and a box (12 quads); 200 IcoSphere(r, 3)
    with r ~ U[0.05, 0.3] inside the room; the jittered subdiv-5 icosphere (radius 1, sigma 0.01).
    Sphere s uses uniform draws 10 s .. 10 s + 9 of stream seed + 3 (x, y, z, r, then 6 for its material);
    the big icosphere and the box take the next 12."""
    import vierkant_amd as vk
    b = _SceneBuilder()
    white = b.material(vk.material((0.8, 0.8, 0.8), roughness=1.0))
    red = b.material(vk.material((0.8, 0.1, 0.1), roughness=1.0))
    green = b.material(vk.material((0.1, 0.8, 0.1), roughness=1.0))
    light = b.material(vk.material((0.0, 0.0, 0.0), emission=(1.0, 0.9, 0.8), emission_strength=8.0))
    room = _box_quads((-3, 0, -3), (3, 6, 3), inward=True)
    b.quads([room[0], room[1], room[2]], white)       # floor, ceiling, back
    b.quads([room[4]], red)
    b.quads([room[5]], green)
    b.quads([([(-1, 5.99, -1), (1, 5.99, -1), (1, 5.99, 1), (-1, 5.99, 1)], (0, -1, 0))], light)
    u = uniform(seed + 3, 0, 10 * n_spheres + 12).reshape(-1).astype(np.float64)
    for s in range(n_spheres):
        x = u[10 * s: 10 * s + 10]
        r = 0.05 + 0.25 * x[3]
        c = np.array([-2.7 + 5.4 * x[0], 0.3 + 4.2 * x[1], -2.7 + 5.4 * x[2]])
        p, i = vk.icosphere(float(r), 3)
        n = p.astype(np.float64)
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        b.add((p + c).astype(np.float32), n.astype(np.float32), i, b.material(_random_material(x[4:10])))
    x = u[10 * n_spheres:]
    p0, i = vk.icosphere(1.0, 5)
    n = p0.astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    pj, _ = jittered_icosphere(5, seed=seed)
    b.add((pj + np.array([-1.2, 1.2, 0.3], np.float32)).astype(np.float32), n.astype(np.float32), i,
          b.material(_random_material(x[0:6])))
    b.quads(_box_quads((0.4, 0.0, -1.8), (1.8, 1.4, -0.4), inward=False), b.material(_random_material(x[6:12])))
    return b.scene(environment=1.0)


def room_camera():
    """eye (0, 3, 10.5) looking down -z (slightly down), 45 degree vertical field of view"""
    import vierkant_amd as vk
    return vk.make_camera(eye=(0.0, 3.0, 10.5), forward=(0.0, -0.05, -1.0), up=(0.0, 1.0, 0.0),
                          fov_y=float(np.radians(45.0)))


def room_lights():
    """analytic lights for the room scene, one of each kind the tracer shades: a rect panel and a sphere near
    the ceiling, a spot aimed at the back wall, a tube along the back wall, an omni near the floor and a sun
    disc shining in through the open front"""
    import vierkant_amd as vk

    def unit(v):
        v = np.asarray(v, np.float64)
        return tuple((v / np.linalg.norm(v)).astype(np.float32))

    return np.array([
        vk.light(vk.LIGHT_RECT, position=(1.5, 5.9, 0.0), direction=(0.0, -1.0, 0.0), tangent=(1.0, 0.0, 0.0),
                 color=(1.0, 0.95, 0.9), intensity=12.0, size=(0.5, 0.4)),
        vk.light(vk.LIGHT_SPHERE, position=(-1.5, 4.5, 1.0), color=(1.0, 0.8, 0.6), intensity=30.0, size=(0.2, 0.0)),
        vk.light(vk.LIGHT_SPOT, position=(0.0, 5.5, 2.5), direction=unit((0.0, -0.6, -1.0)), color=(0.6, 0.8, 1.0),
                 intensity=40.0, inner_cone_angle=0.2, outer_cone_angle=0.5),
        vk.light(vk.LIGHT_TUBE, position=(0.0, 3.0, -2.8), direction=(1.0, 0.0, 0.0), color=(0.9, 1.0, 0.9),
                 intensity=20.0, size=(0.05, 1.0)),
        vk.light(vk.LIGHT_OMNI, position=(2.0, 1.0, 2.0), intensity=5.0),
        vk.light(vk.LIGHT_DIRECTIONAL, direction=unit((0.2, -0.5, -0.85)), intensity=3.0, angular_size=0.05),
    ], dtype=vk.LIGHT_DTYPE)


def object_meshes(scene):
    """a scene's triangles split into meshes at every change of object (each object's triangles are
    contiguous), with one identity instance per mesh in order: (mesh_first, PT_INSTANCE_DTYPE records) for
    PathTracer.set_instanced_scene, which then renders the flat scene's image"""
    import vierkant_amd as vk
    to = np.asarray(scene["tri_object"])
    cuts = np.flatnonzero(np.diff(to)) + 1 if len(to) else np.zeros(0, np.int64)
    mf = np.concatenate([[0], cuts, [len(to)]]).astype(np.uint32)
    return mf, vk.pt_instances(np.arange(len(mf) - 1))


def _unit_quaternions(u):
    """uniform random rotations (x, y, z, w) from 3 uniforms per row (Shoemake's subgroup method)"""
    u = np.asarray(u, np.float64).reshape(-1, 3)
    a, b = np.sqrt(1.0 - u[:, 0]), np.sqrt(u[:, 0])
    t1, t2 = 2.0 * np.pi * u[:, 1], 2.0 * np.pi * u[:, 2]
    return np.column_stack([a * np.sin(t1), a * np.cos(t1), b * np.sin(t2), b * np.cos(t2)]).astype(np.float32)


def room_instanced(n_spheres=200, n_sphere_meshes=8, seed=0):
    """The room as a two-level scene. Mesh 0: the room's quads, emitter and box (one identity instance).
    Meshes 1..n_sphere_meshes: a unit IcoSphere(1, 3), each with its own material, shared by n_spheres
    instances (sphere s uses mesh 1 + s % n_sphere_meshes) with uniform scale r ~ U[0.05, 0.3], a uniformly
    random rotation and a centre inside the room. Last mesh: the jittered subdiv-5 icosphere around the
    origin, placed at (-1.2, 1.2, 0.3) with a random rotation and non-uniform scale (1.0, 0.8, 1.1).
    Uniform draws: stream seed + 5, 7 per sphere (centre, radius, rotation), then 3 + 3 x 6 material draws.
    Returns (scene, mesh_first, instances)."""
    import vierkant_amd as vk
    b = _SceneBuilder()
    u = uniform(seed + 5, 0, 7 * n_spheres + 3 + 6 * (n_sphere_meshes + 2)).reshape(-1).astype(np.float64)
    white = b.material(vk.material((0.8, 0.8, 0.8), roughness=1.0))
    red = b.material(vk.material((0.8, 0.1, 0.1), roughness=1.0))
    green = b.material(vk.material((0.1, 0.8, 0.1), roughness=1.0))
    light = b.material(vk.material((0.0, 0.0, 0.0), emission=(1.0, 0.9, 0.8), emission_strength=8.0))
    room = _box_quads((-3, 0, -3), (3, 6, 3), inward=True)
    b.quads([room[0], room[1], room[2]], white)
    b.quads([room[4]], red)
    b.quads([room[5]], green)
    b.quads([([(-1, 5.99, -1), (1, 5.99, -1), (1, 5.99, 1), (-1, 5.99, 1)], (0, -1, 0))], light)
    m0 = 7 * n_spheres + 3
    b.quads(_box_quads((0.4, 0.0, -1.8), (1.8, 1.4, -0.4), inward=False),
            b.material(_random_material(u[m0:m0 + 6])))
    first = [0, sum(len(i) // 3 for i in b.idx)]
    p, i = vk.icosphere(1.0, 3)
    n = p.astype(np.float64)
    n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    for k in range(n_sphere_meshes):
        mo = m0 + 6 * (k + 1)
        b.add(p, n, i, b.material(_random_material(u[mo:mo + 6])))
        first.append(first[-1] + len(i) // 3)
    pb, ib = vk.icosphere(1.0, 5)
    nb = pb.astype(np.float64)
    nb = (nb / np.linalg.norm(nb, axis=1, keepdims=True)).astype(np.float32)
    pj, _ = jittered_icosphere(5, seed=seed)
    mo = m0 + 6 * (n_sphere_meshes + 1)
    b.add(pj, nb, ib, b.material(_random_material(u[mo:mo + 6])))
    first.append(first[-1] + len(ib) // 3)
    x = u[:7 * n_spheres].reshape(n_spheres, 7)
    centres = np.column_stack([-2.7 + 5.4 * x[:, 0], 0.3 + 4.2 * x[:, 1], -2.7 + 5.4 * x[:, 2]])
    radii = 0.05 + 0.25 * x[:, 3]
    meshes = np.concatenate([[0], 1 + np.arange(n_spheres) % n_sphere_meshes, [n_sphere_meshes + 1]])
    trans = np.vstack([np.zeros((1, 3)), centres, [[-1.2, 1.2, 0.3]]])
    rots = np.vstack([[[0.0, 0.0, 0.0, 1.0]], _unit_quaternions(x[:, 4:7]),
                      _unit_quaternions(u[7 * n_spheres:7 * n_spheres + 3])])
    scales = np.vstack([np.ones((1, 3)), np.repeat(radii[:, None], 3, axis=1), [[1.0, 0.8, 1.1]]])
    inst = vk.pt_instances(meshes, trans, rots, scales)
    return b.scene(environment=1.0), np.asarray(first, np.uint32), inst


def flatten_instances(scene, mesh_first, instances):
    """the two-level scene with every placement's triangles moved into world space (translation +
    rotate(q, scale * p), evaluated in double; normals turned by the rotation only, as the shading turns
    them), one object per placed mesh object: the same geometry as a flat scene"""
    pos, nrm, idx, tob, omat = [], [], [], [], []
    nv = 0
    for rec in instances:
        a, b = int(mesh_first[rec["mesh"]]), int(mesh_first[rec["mesh"] + 1])
        tri = np.asarray(scene["indices"]).reshape(-1, 3)[a:b]
        used, inv = np.unique(tri.reshape(-1), return_inverse=True)
        x, y, z, w = (float(c) for c in rec["rotation"])
        r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        p = np.asarray(scene["positions"])[used].astype(np.float64) * rec["scale"].astype(np.float64)
        pos.append((p @ r.T + rec["translation"].astype(np.float64)).astype(np.float32))
        nrm.append((np.asarray(scene["normals"])[used].astype(np.float64) @ r.T).astype(np.float32))
        idx.append(inv.astype(np.uint32) + nv)
        nv += len(used)
        uo, oinv = np.unique(np.asarray(scene["tri_object"])[a:b], return_inverse=True)
        tob.append((oinv + len(omat)).astype(np.uint32))
        omat.extend(np.asarray(scene["object_material"])[uo])
    return {"positions": np.concatenate(pos), "normals": np.concatenate(nrm), "indices": np.concatenate(idx),
            "tri_object": np.concatenate(tob), "object_material": np.asarray(omat, np.uint32),
            "materials": scene["materials"], "environment": scene.get("environment", 1.0)}


def camera_basis(camera, width, height):
    """the tracer's camera basis, computed like the library does (double precision, rounded once):
    returns (eye, forward, right, up, tan_half_x, tan_half_y) as float32"""
    import math
    fw0 = [float(camera.forward[k]) for k in range(3)]
    up0 = [float(camera.up[k]) for k in range(3)]

    def nrm(v):
        l = math.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
        return [v[0] / l, v[1] / l, v[2] / l]

    def crs(a, b):
        return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]

    fw = nrm(fw0)
    rt = nrm(crs(fw, up0))
    up = crs(rt, fw)
    th = math.tan(0.5 * float(camera.fov_y))
    f32 = lambda v: np.asarray(v, np.float32)
    return (f32([camera.eye[k] for k in range(3)]), f32(fw), f32(rt), f32(up),
            np.float32(th * width / height), np.float32(th))


def minstd_rand0(seed, n):
    """std::default_random_engine (libstdc++: minstd_rand0) outputs: the per-batch random seeds the
    reference path tracer draws (src/PBRPathTracer.cpp:32, :588)"""
    x = seed % 2147483647
    if x == 0:
        x = 1
    out = []
    for _ in range(n):
        x = (16807 * x) % 2147483647
        out.append(x)
    return out


# ------------------------------------------------------------------------------------------------
# configuration 4: many independent scenes for the batched launch
# ------------------------------------------------------------------------------------------------
def many_scenes(n_scenes=256, seed=0, tri_range=(1000, 500000), ray_range=(16384, 2097152)):
    """Yields (positions, indices, rays) for `n_scenes` scenes. Scene s draws (from
    numpy.random.default_rng(seed), in order) its triangle count T and ray count R log-uniform in the given
    ranges; it is a 4 x 4 fBm heightfield patch with about T/2 triangles plus round(T / 2560) IcoSpheres of
    subdivision 3 (scaled copies of the unit one, radius U[0.05, 0.3], each vertex copy jittered by
    N(0, 0.01 r)), and R rays with origins uniform in 1.5x the scene box and directions uniform on S^2."""
    rng = np.random.default_rng(seed)
    unit_pos, unit_idx = vk.icosphere(1.0, 3)
    lt, lr = np.log(tri_range), np.log(ray_range)
    for s in range(n_scenes):
        n_tri = int(np.exp(rng.uniform(lt[0], lt[1])))
        n_ray = int(np.exp(rng.uniform(lr[0], lr[1])))
        g = max(1, int(round(np.sqrt(n_tri / 4.0))))
        pos, idx = vk.plane(4.0, 4.0, g, g)
        pos[:, 1] = (0.2 * fbm(pos[:, 0], pos[:, 2], seed=s)).astype(np.float32)
        k = max(1, int(round(n_tri / 2560.0)))
        r = rng.uniform(0.05, 0.3, k)
        c = np.stack([rng.uniform(-2, 2, k), rng.uniform(0.1, 1.0, k), rng.uniform(-2, 2, k)], axis=1)
        sp = unit_pos[None, :, :] * r[:, None, None] + c[:, None, :]
        sp = sp + rng.normal(0.0, 1.0, sp.shape) * (0.01 * r)[:, None, None]
        base = len(pos) + np.arange(k, dtype=np.uint32)[:, None] * np.uint32(len(unit_pos))
        si = (unit_idx[None, :] + base).reshape(-1)
        positions = np.ascontiguousarray(np.concatenate([pos, sp.reshape(-1, 3).astype(np.float32)]))
        indices = np.ascontiguousarray(np.concatenate([idx, si.astype(np.uint32)]))
        lo, hi = positions.min(0), positions.max(0)
        center, half = 0.5 * (lo + hi), 0.75 * (hi - lo)
        o = (center + (rng.uniform(-1.0, 1.0, (n_ray, 3)) * half)).astype(np.float32)
        z = rng.uniform(-1.0, 1.0, n_ray)
        phi = rng.uniform(0.0, 2.0 * np.pi, n_ray)
        rxy = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        d = glm_normalize(np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], axis=1).astype(np.float32))
        yield positions, indices, make_rays(o, d)
