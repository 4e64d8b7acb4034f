"""Radiance cache, CPU restatement (numpy) - TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, as the
reference the GPU radiance cache is compared with; no product path calls it.

The reference has no radiance cache: this component is new. It follows the published methods it is
built from, each restated here from the mathematics:
  - multiresolution hash encoding (Mueller, Evans, Schied, Keller 2022, "Instant Neural Graphics
    Primitives", sec. 3): L levels with resolutions floor(N_min * b^l), b = exp((ln N_max - ln N_min) /
    (L - 1)); a level whose (N + 1)^3 corners fit the table is indexed densely, finer levels by the
    spatial hash (x * 1) xor (y * 2654435761) xor (z * 805459861) mod T; trilinear interpolation of the
    8 corners' F = 2 features;
  - real spherical harmonics up to degree 3 (16 coefficients) of the outgoing direction (the standard
    orthonormal real basis; the constants are sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) factors);
  - neural radiance caching (Mueller, Rousselle, Novak, Keller 2021): a small ReLU MLP without biases,
    relative L2 loss (y - t)^2 / (sg(mean y)^2 + 0.01), online training with Adam;
  - Adam (Kingma & Ba 2015) with bias correction; a parameter whose gradient is exactly zero this step
    is left unchanged (sparse update of untouched hash-table entries, as in instant-ngp).

Network: input 64 -> 5 hidden layers of 64 (ReLU) -> 16 outputs (3 used: rgb radiance), no biases;
an input fixed at 1.0 gives the first layer its bias. Matrix inputs are rounded to bfloat16 (round to
nearest even) and products accumulate in fp32 or wider, as on the matrix cores.

Record (16 floats): pos.xyz, roughness | wo.xyz, metalness | normal.xyz, transmission | albedo.rgb, -.

Logical input order (64): for g in 0..3: [16g + 0..7] = hash levels 4g..4g+3 (2 features each),
[16g + 8..11] = SH coefficients 4g..4g+3, [16g + 12..15] = extras 4g..4g+3 with
extras = (nx, ny, nz, roughness, albedo rgb, metalness, transmission, 1, 0, 0, 0, 0, 0, 0).
"""
import math

import numpy as np

N_LEVELS = 16
N_FEATURES = 2
LOG2_T = 19
N_MIN = 16
N_MAX = 2048
WIDTH = 64
N_HIDDEN = 5
N_OUT = 16
LAYER_SHAPES = [(WIDTH, WIDTH)] * N_HIDDEN + [(N_OUT, WIDTH)]      # (out, in), W1..W6
N_MLP = sum(o * i for o, i in LAYER_SHAPES)                         # 21,504
PRIMES = (1, 2654435761, 805459861)
LOSS_EPS = 0.01


def resolutions(n_levels=N_LEVELS, n_min=N_MIN, n_max=N_MAX):
    b = math.exp((math.log(n_max) - math.log(n_min)) / (n_levels - 1))
    return [int(math.floor(n_min * b ** l)) for l in range(n_levels)]


def bf16(x):
    """round fp32 to bfloat16 (nearest even), returned as fp32"""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).reshape(x.shape)
    nan = np.isnan(x)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


def sh16(d):
    """real spherical harmonics, degrees 0..3, of unit directions d (N, 3) -> (N, 16) float32"""
    d = d.astype(np.float32)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz = x * x, y * y, z * z
    out = np.empty((len(d), 16), np.float32)
    out[:, 0] = np.float32(0.28209479177387814)
    out[:, 1] = np.float32(-0.48860251190291987) * y
    out[:, 2] = np.float32(0.48860251190291987) * z
    out[:, 3] = np.float32(-0.48860251190291987) * x
    out[:, 4] = np.float32(1.0925484305920792) * (x * y)
    out[:, 5] = np.float32(-1.0925484305920792) * (y * z)
    out[:, 6] = np.float32(0.94617469575755997) * zz - np.float32(0.31539156525251999)
    out[:, 7] = np.float32(-1.0925484305920792) * (x * z)
    out[:, 8] = np.float32(0.54627421529603959) * (xx - yy)
    out[:, 9] = np.float32(0.59004358992664352) * (y * (np.float32(3.0) * xx - yy))
    out[:, 10] = np.float32(2.8906114426405538) * ((x * y) * z)
    out[:, 11] = np.float32(0.45704579946446572) * (y * (np.float32(1.0) - np.float32(5.0) * zz))
    out[:, 12] = np.float32(0.3731763325901154) * (z * (np.float32(5.0) * zz - np.float32(3.0)))
    out[:, 13] = np.float32(0.45704579946446572) * (x * (np.float32(1.0) - np.float32(5.0) * zz))
    out[:, 14] = np.float32(1.4453057213202769) * (z * (xx - yy))
    out[:, 15] = np.float32(0.59004358992664352) * (x * (xx - np.float32(3.0) * yy))
    return out


def grid_corners(records, bmin, inv_extent, log2_t=LOG2_T, res=None):
    """per level: (N, 8) table indices and (N, 8) trilinear weights; corner k = (k & 1, k >> 1 & 1, k >> 2)"""
    res = res or resolutions()
    t = 1 << log2_t
    p = (records[:, 0:3].astype(np.float32) - np.asarray(bmin, np.float32)) * np.asarray(inv_extent, np.float32)
    p = np.clip(p, np.float32(0.0), np.float32(1.0))
    idx, wts = [], []
    for n in res:
        x = p * np.float32(n)
        c0 = np.minimum(np.floor(x), np.float32(n - 1)).astype(np.int64)
        f = (x - c0.astype(np.float32)).astype(np.float32)
        dense = (n + 1) ** 3 <= t
        li = np.empty((len(p), 8), np.int64)
        lw = np.empty((len(p), 8), np.float32)
        for k in range(8):
            d = np.array([k & 1, (k >> 1) & 1, k >> 2])
            c = c0 + d
            w = np.ones(len(p), np.float32)
            for a in range(3):
                w = w * (f[:, a] if d[a] else (np.float32(1.0) - f[:, a]))
            if dense:
                li[:, k] = c[:, 0] + (n + 1) * (c[:, 1] + (n + 1) * c[:, 2])
            else:
                h = (c[:, 0] * PRIMES[0]) ^ (c[:, 1] * PRIMES[1]) ^ (c[:, 2] * PRIMES[2])
                li[:, k] = (h & 0xFFFFFFFF) & (t - 1)
            lw[:, k] = w
        idx.append(li)
        wts.append(lw)
    return idx, wts


def encode(records, grid, bmin, inv_extent):
    """records (N, 16), grid (L, T, 2) float32 (values as stored in fp16) -> x0 (N, 64) float32"""
    n_levels = grid.shape[0]
    idx, wts = grid_corners(records, bmin, inv_extent, int(round(math.log2(grid.shape[1]))))
    feats = np.zeros((len(records), n_levels, 2), np.float32)
    for l in range(n_levels):
        for k in range(8):
            feats[:, l, :] = feats[:, l, :] + wts[l][:, k:k + 1] * grid[l][idx[l][:, k]]
    sh = sh16(records[:, 4:7])
    extras = np.zeros((len(records), 16), np.float32)
    extras[:, 0:3] = records[:, 8:11]
    extras[:, 3] = records[:, 3]
    extras[:, 4:7] = records[:, 12:15]
    extras[:, 7] = records[:, 7]
    extras[:, 8] = records[:, 11]
    extras[:, 9] = 1.0
    x0 = np.zeros((len(records), 64), np.float32)
    for g in range(4):
        x0[:, 16 * g: 16 * g + 8] = feats[:, 4 * g: 4 * g + 4, :].reshape(-1, 8)
        x0[:, 16 * g + 8: 16 * g + 12] = sh[:, 4 * g: 4 * g + 4]
        x0[:, 16 * g + 12: 16 * g + 16] = extras[:, 4 * g: 4 * g + 4]
    return x0


def hash_slot(level, feature):
    """logical input index of a level's feature"""
    return 16 * (level // 4) + 2 * (level % 4) + feature


def split_weights(flat):
    out, o = [], 0
    for r, c in LAYER_SHAPES:
        out.append(np.asarray(flat[o: o + r * c], np.float32).reshape(r, c))
        o += r * c
    return out


def _identity(x):
    return np.asarray(x)


def forward(x0, weights, rnd=bf16, dtype=np.float32):
    """bf16 matrix inputs, float64 accumulation rounded to fp32 per layer; returns (y, activations, z).
    rnd=_identity, dtype=np.float64 evaluates the same network without rounding (gradient checks)."""
    acts = [x0.astype(dtype)]
    zs = []
    a = acts[0]
    for i, w in enumerate(weights):
        z = (rnd(a).astype(np.float64) @ rnd(w).astype(np.float64).T).astype(dtype)
        zs.append(z)
        if i < len(weights) - 1:
            a = np.maximum(z, 0.0).astype(dtype)
            acts.append(a)
    return zs[-1], acts, zs


def loss_grad(y, targets, n_total, dtype=np.float32):
    """relative L2 over rgb: sum_c (y_c - t_c)^2 / (sg(mean(y))^2 + 0.01) / n_total; returns (loss, dy).
    The denominator is a constant for differentiation (stop-gradient)."""
    y3 = y[:, :3].astype(dtype)
    t3 = targets[:, :3].astype(dtype)
    ybar = (y3[:, 0] + y3[:, 1] + y3[:, 2]) / dtype(3.0)
    denom = ybar * ybar + dtype(LOSS_EPS)
    diff = y3 - t3
    loss = float(np.sum((diff * diff) / denom[:, None]) / n_total)
    dy = np.zeros(y.shape, dtype=dtype)
    dy[:, :3] = (dtype(2.0) * diff / denom[:, None]) / dtype(n_total)
    return loss, dy


def backward(dy, acts, zs, weights, rnd=bf16, dtype=np.float32):
    """gradients of every weight matrix and of the input features; the matrix inputs (activations and
    back-propagated gradients) are rounded to bf16 like the forward pass"""
    dws = [None] * len(weights)
    dz = dy.astype(dtype)
    for i in range(len(weights) - 1, -1, -1):
        dws[i] = (rnd(dz).astype(np.float64).T @ rnd(acts[i]).astype(np.float64)).astype(dtype)
        dh = (rnd(dz).astype(np.float64) @ rnd(weights[i]).astype(np.float64)).astype(dtype)
        if i > 0:
            dz = np.where(zs[i - 1] > 0, dh, 0.0).astype(dtype)
        else:
            dx0 = dh
    return dws, dx0


def grid_grad(records, dx0, grid_shape, bmin, inv_extent):
    g = np.zeros(grid_shape, np.float64)
    idx, wts = grid_corners(records, bmin, inv_extent, int(round(math.log2(grid_shape[1]))))
    for l in range(grid_shape[0]):
        for c in range(2):
            d = dx0[:, hash_slot(l, c)].astype(np.float64)
            for k in range(8):
                np.add.at(g[l, :, c], idx[l][:, k], wts[l][:, k].astype(np.float64) * d)
    return g.astype(np.float32)


def adam(param, grad, m, v, step, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-15):
    """one Adam step (in place on copies); a parameter with grad == 0 is left unchanged"""
    param, m, v = param.copy(), m.copy(), v.copy()
    g = grad.astype(np.float32)
    touched = g != 0
    b1, b2 = np.float32(beta1), np.float32(beta2)
    m_new = b1 * m + (np.float32(1.0) - b1) * g
    v_new = b2 * v + (np.float32(1.0) - b2) * (g * g)
    c1 = np.float32(1.0 / (1.0 - beta1 ** step))
    c2 = np.float32(1.0 / (1.0 - beta2 ** step))
    upd = np.float32(lr) * (m_new * c1) / (np.sqrt(v_new * c2) + np.float32(eps))
    param[touched] = (param - upd)[touched]
    m[touched] = m_new[touched]
    v[touched] = v_new[touched]
    return param, m, v


def query(records, grid, mlp_flat, bmin, inv_extent):
    """cache prediction: max(y_rgb, 0) with the stored (fp16) grid"""
    x0 = encode(records, grid, bmin, inv_extent)
    y, _, _ = forward(x0, split_weights(mlp_flat))
    return np.maximum(y[:, :3], np.float32(0.0))


def train_gradients(records, targets, grid, mlp_flat, bmin, inv_extent):
    """loss, flat MLP gradient (21,504) and grid gradient (L, T, 2) of one training batch"""
    weights = split_weights(mlp_flat)
    x0 = encode(records, grid, bmin, inv_extent)
    y, acts, zs = forward(x0, weights)
    loss, dy = loss_grad(y, targets, len(records))
    dws, dx0 = backward(dy, acts, zs, weights)
    g_grid = grid_grad(records, dx0, grid.shape, bmin, inv_extent)
    return loss, np.concatenate([d.reshape(-1) for d in dws]), g_grid
