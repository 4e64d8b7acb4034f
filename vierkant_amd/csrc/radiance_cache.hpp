// radiance_cache.hpp - neural radiance cache: multiresolution hash grid + fully fused 64-wide ReLU MLP on
// gfx950 BF16 matrix cores. Shared by the cache's own entry points and the path tracer's cache mode.
//
// Methods: hash encoding of Mueller et al. 2022 (Instant NGP, sec. 3), neural radiance caching of
// Mueller et al. 2021 (relative L2 loss, online Adam training), real spherical harmonics of the outgoing
// direction. The reference renderer has no such cache; nothing here follows a reference file.
//
// Matrix-core mapping (v_mfma_f32_16x16x32_bf16, wave64). The network is evaluated transposed: one
// 16-sample tile is the N dimension and features are M/K, so a lane keeps one sample across all layers:
//   A (weights)      lane l holds A[l % 16][8 (l / 16) + j], j = 0..7 (16 B per lane)
//   B (activations)  lane l holds B[8 (l / 16) + j][l % 16]
//   D (fp32)         lane l holds D[4 (l / 16) + i][l % 16], i = 0..3
// The D registers of one layer are the next layer's B operand after a fixed permutation of the K index
// (feature phi(s, g, j) = 16 (2 s + j / 4) + 4 g + j % 4 at k-step s, lane group g, element j); the
// weights are stored pre-permuted ("swizzled") so no data moves between lanes. Weight gradients contract
// over samples and read both operands from LDS in [feature][sample] layout.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "context.hpp"

namespace vkrt::rc
{

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));

constexpr uint32_t LEVELS = 16;        // 4 lane groups x 4 levels
constexpr uint32_t WIDTH = 64;
constexpr uint32_t OUTPUTS = 16;       // rgb + 13 padding outputs
constexpr uint32_t N_HIDDEN = 5;
constexpr uint32_t N_LAYERS = 6;       // W1..W5 (64 x 64), W6 (16 x 64)
constexpr uint32_t N_MLP = N_HIDDEN * WIDTH * WIDTH + OUTPUTS * WIDTH;// 21,504
constexpr uint32_t LAYER_ELEMS = WIDTH * WIDTH;                       // swizzled bf16 elements per layer
constexpr uint32_t RECORD_F4 = 4;      // record: pos+roughness, wo+metalness, normal+transmission, albedo
constexpr float LOSS_EPS = 0.01f;

//! grid geometry passed by value to kernels
struct grid_desc
{
    uint32_t res[LEVELS];
    uint32_t dense_mask;// bit l set: level l is indexed densely
    uint32_t log2_t;
    float bmin[3];
    float inv_extent[3];
};

//! device pointers of one cache
struct cache_dev
{
    const __half2 *grid;  // [LEVELS][T] stored features
    const bf16x8 *w_fwd;  // swizzled forward weights
    const bf16x8 *w_bwd;  // swizzled transposed weights (back-propagation)
    grid_desc g;
};

// ---------------------------------------------------------------------------------------------------
// encoding
// ---------------------------------------------------------------------------------------------------

//! real spherical harmonics of degrees 0..3 at a unit direction
__device__ __forceinline__ void sh16(float x, float y, float z, float out[16])
{
    const float xx = x * x, yy = y * y, zz = z * z;
    out[0] = 0.28209479177387814f;
    out[1] = -0.48860251190291987f * y;
    out[2] = 0.48860251190291987f * z;
    out[3] = -0.48860251190291987f * x;
    out[4] = 1.0925484305920792f * (x * y);
    out[5] = -1.0925484305920792f * (y * z);
    out[6] = 0.94617469575755997f * zz - 0.31539156525251999f;
    out[7] = -1.0925484305920792f * (x * z);
    out[8] = 0.54627421529603959f * (xx - yy);
    out[9] = 0.59004358992664352f * (y * (3.0f * xx - yy));
    out[10] = 2.8906114426405538f * ((x * y) * z);
    out[11] = 0.45704579946446572f * (y * (1.0f - 5.0f * zz));
    out[12] = 0.3731763325901154f * (z * (5.0f * zz - 3.0f));
    out[13] = 0.45704579946446572f * (x * (1.0f - 5.0f * zz));
    out[14] = 1.4453057213202769f * (z * (xx - yy));
    out[15] = 0.59004358992664352f * (x * (xx - 3.0f * yy));
}

//! position normalized to the cache bounds, clamped to [0, 1]
__device__ __forceinline__ void unit_position(const grid_desc &g, const float4 &r0, float p[3])
{
    p[0] = fminf(fmaxf((r0.x - g.bmin[0]) * g.inv_extent[0], 0.0f), 1.0f);
    p[1] = fminf(fmaxf((r0.y - g.bmin[1]) * g.inv_extent[1], 0.0f), 1.0f);
    p[2] = fminf(fmaxf((r0.z - g.bmin[2]) * g.inv_extent[2], 0.0f), 1.0f);
}

//! the 8 corner table indices and trilinear weights of one level; corner k = (k & 1, k >> 1 & 1, k >> 2)
__device__ __forceinline__ void level_corners(const grid_desc &g, uint32_t level, const float p[3], uint32_t idx[8],
                                              float w[8])
{
    const uint32_t n = g.res[level];
    const float nf = static_cast<float>(n);
    uint32_t c0[3];
    float f[3];
    for(int a = 0; a < 3; ++a)
    {
        const float x = p[a] * nf;
        const float fl = fminf(floorf(x), static_cast<float>(n - 1));
        c0[a] = static_cast<uint32_t>(fl);
        f[a] = x - fl;
    }
    const bool dense = (g.dense_mask >> level) & 1u;
    const uint32_t mask = (1u << g.log2_t) - 1u;
#pragma unroll
    for(int k = 0; k < 8; ++k)
    {
        const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        const uint32_t cx = c0[0] + dx, cy = c0[1] + dy, cz = c0[2] + dz;
        const float wx = dx ? f[0] : 1.0f - f[0];
        const float wy = dy ? f[1] : 1.0f - f[1];
        const float wz = dz ? f[2] : 1.0f - f[2];
        w[k] = (wx * wy) * wz;
        idx[k] = dense ? cx + (n + 1u) * (cy + (n + 1u) * cz)
                       : ((cx * 1u) ^ (cy * 2654435761u) ^ (cz * 805459861u)) & mask;
    }
}

//! the 16 logical inputs 16 g .. 16 g + 15 that lane group g feeds to the first layer
__device__ __forceinline__ void encode_group(const cache_dev &c, const float4 rec[4], uint32_t group, float v[16])
{
    float p[3];
    unit_position(c.g, rec[0], p);
    const size_t table = size_t(1) << c.g.log2_t;
#pragma unroll
    for(uint32_t m = 0; m < 4; ++m)
    {
        const uint32_t level = 4 * group + m;
        uint32_t idx[8];
        float w[8];
        level_corners(c.g, level, p, idx, w);
        float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
        for(int k = 0; k < 8; ++k)
        {
            const float2 e = __half22float2(c.grid[level * table + idx[k]]);
            f0 = f0 + w[k] * e.x;
            f1 = f1 + w[k] * e.y;
        }
        v[2 * m] = f0;
        v[2 * m + 1] = f1;
    }
    float sh[16];
    sh16(rec[1].x, rec[1].y, rec[1].z, sh);
    const float extras[16] = {rec[2].x, rec[2].y, rec[2].z, rec[0].w, rec[3].x, rec[3].y, rec[3].z, rec[1].w,
                              rec[2].w, 1.0f,     0.0f,     0.0f,     0.0f,     0.0f,     0.0f,     0.0f};
#pragma unroll
    for(uint32_t k = 0; k < 4; ++k)
    {
        v[8 + k] = sh[4 * group + k];
        v[12 + k] = extras[4 * group + k];
    }
}

__device__ __forceinline__ bf16x8 pack8(const float *v)
{
    bf16x8 r;
#pragma unroll
    for(int j = 0; j < 8; ++j) r[j] = static_cast<__bf16>(v[j]);
    return r;
}

//! D registers (4 row tiles x 4) -> the next layer's two B operands (K permutation phi)
__device__ __forceinline__ void pack_next(const f32x4 d[4], bf16x8 b[2])
{
#pragma unroll
    for(int s = 0; s < 2; ++s)
#pragma unroll
        for(int j = 0; j < 8; ++j) b[s][j] = static_cast<__bf16>(d[2 * s + j / 4][j % 4]);
}

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

//! forward pass of T tiles (16 samples each) held by one wave; returns the output D tile per sample tile
//! (lane group 0 holds outputs 0..3, i.e. rgb + one padding output)
template<int T>
__device__ __forceinline__ void mlp_forward(const bf16x8 *__restrict__ w_fwd, uint32_t lane, bf16x8 b[T][2],
                                            f32x4 y[T])
{
    for(uint32_t l = 0; l < N_HIDDEN; ++l)
    {
        f32x4 acc[T][4];
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            const bf16x8 a0 = w_fwd[(l * LAYER_ELEMS) / 8 + (r * 2 + 0) * 64 + lane];
            const bf16x8 a1 = w_fwd[(l * LAYER_ELEMS) / 8 + (r * 2 + 1) * 64 + lane];
#pragma unroll
            for(int t = 0; t < T; ++t)
            {
                f32x4 z = {0.f, 0.f, 0.f, 0.f};
                z = mfma(a0, b[t][0], z);
                acc[t][r] = mfma(a1, b[t][1], z);
            }
        }
#pragma unroll
        for(int t = 0; t < T; ++t)
        {
#pragma unroll
            for(int r = 0; r < 4; ++r)
#pragma unroll
                for(int i = 0; i < 4; ++i) acc[t][r][i] = fmaxf(acc[t][r][i], 0.0f);
            pack_next(acc[t], b[t]);
        }
    }
    const bf16x8 a0 = w_fwd[(N_HIDDEN * LAYER_ELEMS) / 8 + 0 * 64 + lane];
    const bf16x8 a1 = w_fwd[(N_HIDDEN * LAYER_ELEMS) / 8 + 1 * 64 + lane];
#pragma unroll
    for(int t = 0; t < T; ++t)
    {
        f32x4 z = {0.f, 0.f, 0.f, 0.f};
        z = mfma(a0, b[t][0], z);
        y[t] = mfma(a1, b[t][1], z);
    }
}

// ---------------------------------------------------------------------------------------------------
// weight swizzles (logical W[out][in] -> bf16 element of the A operand arrays)
// ---------------------------------------------------------------------------------------------------

//! forward: A = W_l, rows = outputs, K = inputs in the order the B operand presents them
__host__ __device__ inline uint32_t fwd_pos(uint32_t layer, uint32_t out, uint32_t in)
{
    const uint32_t r = out / 16, row = out % 16;
    uint32_t g, s, j;
    if(layer == 0)
    {
        g = in / 16;
        s = (in % 16) / 8;
        j = in % 8;
    }
    else
    {
        const uint32_t t = in / 16;
        s = t / 2;
        j = 4 * (t % 2) + in % 4;
        g = (in % 16) / 4;
    }
    return layer * LAYER_ELEMS + ((r * 2 + s) * 64 + 16 * g + row) * 8 + j;
}

//! back-propagation: A = W_l^T, rows = inputs, K = outputs as the gradient's B operand presents them
__host__ __device__ inline uint32_t bwd_pos(uint32_t layer, uint32_t out, uint32_t in)
{
    const uint32_t r = in / 16, row = in % 16;
    if(layer < N_HIDDEN)
    {
        const uint32_t t = out / 16;
        const uint32_t s = t / 2, j = 4 * (t % 2) + out % 4, g = (out % 16) / 4;
        return layer * LAYER_ELEMS + ((r * 2 + s) * 64 + 16 * g + row) * 8 + j;
    }
    // W6: 16 outputs, one k-step; lane group g carries outputs 4 g .. 4 g + 3 in elements 0..3
    const uint32_t g = out / 4, j = out % 4;
    return N_HIDDEN * LAYER_ELEMS + (r * 64 + 16 * g + row) * 8 + j;
}

//! offset of W_l's first logical element in the flat [out][in] master array
__host__ __device__ inline uint32_t layer_offset(uint32_t layer) { return layer * LAYER_ELEMS; }

}// namespace vkrt::rc

//! the cache object behind the C ABI
struct vkrt_rc
{
    vkrt_ctx *ctx = nullptr;
    vkrt_rc_params params{};
    vkrt::rc::grid_desc desc{};
    uint64_t n_grid = 0;          // LEVELS * T * 2 parameters
    uint32_t step = 0;            // Adam steps taken
    float grad_scale = 1.0f;      // loss scale of the last step's fp16 grid gradient
    vkrt::device_buffer grid_master, grid_half, grid_grad, grid_grad_f32, grid_m, grid_v;
    // grid_grad holds zeros: the last training step's Adam pass cleared every entry it read as non-zero
    bool grid_grad_clear = false;
    vkrt::device_buffer mlp_master, mlp_grad, mlp_m, mlp_v, w_fwd, w_bwd, partials, loss;
};

namespace vkrt::rc
{
inline cache_dev device_view(const vkrt_rc *c)
{
    return cache_dev{c->grid_half.as<const __half2>(), c->w_fwd.as<const bf16x8>(), c->w_bwd.as<const bf16x8>(),
                     c->desc};
}

//! one training step over `n` device records (RECORD_F4 float4 each, `stride_f4` apart) with targets
//! (float4 rgb, -): gradients, then Adam unless grad_only. With n_dev the count is min(*n_dev, n), read on
//! the device (no host synchronization). Launched on `stream` (the caller orders it against queries).
void train(vkrt_rc *c, const float4 *records, uint32_t stride_f4, const float4 *targets, uint32_t n, bool grad_only,
           const uint32_t *n_dev, hipStream_t stream);

//! predictions max(y_rgb, 0): written to out[i] (float4), or, when pixel_acc is set, added as
//! beta * prediction to pixel_acc[pixel] with pixel and beta taken from the record's fifth float4
void query(vkrt_rc *c, const float4 *records, uint32_t stride_f4, uint32_t n, float4 *out, float4 *pixel_acc,
           const uint32_t *count_on_device);
}// namespace vkrt::rc
