"""adversarial (triangle, origin, direction) cases for the Moller-Trumbore bounds; shared by the host-header
and device parity tests"""
import numpy as np


def adversarial_mt_cases():
    """hits exactly on edges and vertices (u = 0, v = 0, u + v = 1), determinants straddling +-1e-9, tiny and
    huge triangles (determinants beyond 2^100 switch the division-free bounds off), non-finite inputs"""
    rng = np.random.default_rng(11)
    cases = []
    for k in range(400):
        v0, e1, e2 = rng.normal(size=(3, 3)).astype(np.float32)
        s = np.float32(rng.uniform(0, 1))
        which = k % 4
        target = [v0 + s * e1, v0 + s * e2, v0 + s * e1 + (1 - s) * e2, v0 + e1 * (k % 2)][which]   # v=0, u=0, u+v=1, vertex
        d = rng.normal(size=3).astype(np.float32)
        d /= np.linalg.norm(d)
        cases.append((np.concatenate([v0, v0 + e1, v0 + e2]), (target - 2 * d).astype(np.float32), d))
    for k in range(200):                       # rays almost in the triangle's plane: det near +-1e-9
        v0 = np.zeros(3, np.float32)
        e1 = np.array([1, 0, 0], np.float32)
        e2 = np.array([0, 1, 0], np.float32)
        dz = np.float32([1e-9, -1e-9, 1.0001e-9, 0.9999e-9, 2e-9, 5e-10][k % 6]) * np.float32(rng.choice([1, -1]))
        d = np.array([0.3, 0.2, dz], np.float32)
        o = np.array([0.1, 0.1, -dz * 3], np.float32) - 3 * d
        cases.append((np.concatenate([v0, e1, e2]), o.astype(np.float32), d))
    for scale in (1e-20, 1e-10, 1e20, 1e25, 1e30, 3e38):   # tiny and huge triangles
        for k in range(30):
            with np.errstate(over="ignore"):           # 3e38 x normal overflows to +-inf in float32
                tri = (rng.normal(size=9) * scale).astype(np.float32)
                o = (rng.normal(size=3) * scale).astype(np.float32)
            d = rng.normal(size=3).astype(np.float32)
            cases.append((tri, o, d))
    special = [np.inf, -np.inf, np.nan, 0.0]
    for k in range(64):                       # non-finite inputs
        tri = rng.normal(size=9).astype(np.float32)
        o = rng.normal(size=3).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        (tri if k % 2 else d)[k % 3] = special[k % 4]
        cases.append((tri, o, d))
    return [tuple(np.ascontiguousarray(x, np.float32) for x in c) for c in cases]
