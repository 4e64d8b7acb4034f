// radiance_cache.hip - kernels and C ABI of the neural radiance cache (layout and methods: radiance_cache.hpp).
//
//   k_rc_query    per wave 4 x 16 samples: hash-grid + SH encoding in registers, 6 MFMA layers, output
//   k_rc_train    per block 64 samples per iteration: forward (activations to LDS), relative L2 loss,
//                 back-propagation, weight gradients contracted over the samples in LDS, hash-grid gradient
//                 scattered with packed fp16x2 atomics (one per corner, loss-scaled by a power of two);
//                 per-block weight gradients written to a partial buffer
//   k_rc_adam_mlp per weight: sum of the block partials in block order (deterministic), Adam, bf16 swizzles
//   k_rc_adam_grid per grid parameter: Adam where the gradient is non-zero, fp16 copy for the encoder
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "radiance_cache.hpp"

namespace vkrt::rc
{

constexpr int QUERY_BLOCK = 256;  // 4 waves x 64 samples
constexpr int QUERY_TILES = 4;
constexpr int TRAIN_BLOCK = 256;  // 4 waves x 16 samples per iteration
constexpr int TRAIN_SAMPLES = 64;
constexpr int LDS_ROW = 72;       // bf16 per feature row: 64 samples + 8 padding (144 B, 16 B aligned)

__device__ __forceinline__ void load_record(const float4 *records, uint32_t stride_f4, uint32_t i, float4 rec[4])
{
    const float4 *r = records + (size_t)i * stride_f4;
    rec[0] = r[0];
    rec[1] = r[1];
    rec[2] = r[2];
    rec[3] = r[3];
}

__global__ void __launch_bounds__(QUERY_BLOCK) k_rc_query(cache_dev c, const float4 *__restrict__ records,
                                                          uint32_t stride_f4, uint32_t n_host,
                                                          const uint32_t *__restrict__ n_dev, float4 *__restrict__ out,
                                                          float4 *__restrict__ pixel_acc)
{
    const uint32_t n = n_dev ? *n_dev : n_host;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t base = (blockIdx.x * (QUERY_BLOCK / 64) + wave) * (16 * QUERY_TILES);
    if(base >= n) return;// whole wave idle; the MLP below needs all 64 lanes
    const uint32_t group = lane >> 4;
    bf16x8 b[QUERY_TILES][2];
#pragma unroll
    for(int t = 0; t < QUERY_TILES; ++t)
    {
        const uint32_t i = base + 16 * t + (lane & 15u);
        float v[16];
        if(i < n)
        {
            float4 rec[4];
            load_record(records, stride_f4, i, rec);
            encode_group(c, rec, group, v);
        }
        else
            for(int k = 0; k < 16; ++k) v[k] = 0.0f;
        b[t][0] = pack8(v);
        b[t][1] = pack8(v + 8);
    }
    f32x4 y[QUERY_TILES];
    mlp_forward<QUERY_TILES>(c.w_fwd, lane, b, y);
    if(group != 0) return;
#pragma unroll
    for(int t = 0; t < QUERY_TILES; ++t)
    {
        const uint32_t i = base + 16 * t + (lane & 15u);
        if(i >= n) continue;
        const float r = fmaxf(y[t][0], 0.0f), g = fmaxf(y[t][1], 0.0f), bl = fmaxf(y[t][2], 0.0f);
        if(pixel_acc)
        {
            // record's fifth float4: pixel index bits, beta rgb; each pixel has at most one query
            const float4 e = records[(size_t)i * stride_f4 + 4];
            const uint32_t p = __float_as_uint(e.x);
            float4 a = pixel_acc[p];
            a.x = a.x + e.y * r;
            a.y = a.y + e.z * g;
            a.z = a.z + e.w * bl;
            pixel_acc[p] = a;
        }
        else
            out[i] = make_float4(r, g, bl, 0.0f);
    }
}

__global__ void __launch_bounds__(TRAIN_BLOCK) k_rc_train(cache_dev c, const float4 *__restrict__ records,
                                                          uint32_t stride_f4, const float4 *__restrict__ targets,
                                                          uint32_t n_cap, const uint32_t *__restrict__ n_dev,
                                                          half2v *__restrict__ grid_grad, float grad_scale,
                                                          float *__restrict__ partials, float *__restrict__ loss_sum)
{
    // sample count: n_cap, or the device counter clamped to n_cap
    const uint32_t n = n_dev ? min(*n_dev, n_cap) : n_cap;
    __shared__ __bf16 act[N_LAYERS][WIDTH][LDS_ROW];// a_0 = x0, a_i = h_i (i = 1..5), [feature][sample]
    __shared__ __bf16 dzs[WIDTH][LDS_ROW];          // the current layer's output gradient, [feature][sample]
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, group = lane >> 4, col = lane & 15u;
    const uint32_t m_local = 16 * wave + col;
    const size_t table = size_t(1) << c.g.log2_t;

    f32x4 dw[N_HIDDEN][4];// W1..W5: row tile = wave, column tiles 0..3
    f32x4 dw6 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for(int l = 0; l < (int)N_HIDDEN; ++l)
#pragma unroll
        for(int q = 0; q < 4; ++q) dw[l][q] = f32x4{0.f, 0.f, 0.f, 0.f};
    float loss_lane = 0.0f;
    const uint32_t iterations = (n + TRAIN_SAMPLES - 1) / TRAIN_SAMPLES;

    for(uint32_t it = blockIdx.x; it < iterations; it += gridDim.x)
    {
        const uint32_t m = it * TRAIN_SAMPLES + m_local;
        const bool valid = m < n;
        float4 rec[4];
        float v[16];
        if(valid)
        {
            load_record(records, stride_f4, m, rec);
            encode_group(c, rec, group, v);
        }
        else
            for(int k = 0; k < 16; ++k) v[k] = 0.0f;
        __syncthreads();// the previous iteration's readers of act / dzs are done
#pragma unroll
        for(int k = 0; k < 16; ++k) act[0][16 * group + k][m_local] = static_cast<__bf16>(v[k]);
        bf16x8 b[2] = {pack8(v), pack8(v + 8)};

        // ---- forward -------------------------------------------------------------------------------
        uint32_t relu_mask[N_HIDDEN];
#pragma unroll
        for(uint32_t l = 0; l < N_HIDDEN; ++l)
        {
            f32x4 acc[4];
#pragma unroll
            for(int r = 0; r < 4; ++r)
            {
                f32x4 z = {0.f, 0.f, 0.f, 0.f};
                z = mfma(c.w_fwd[(l * LAYER_ELEMS) / 8 + (r * 2 + 0) * 64 + lane], b[0], z);
                acc[r] = mfma(c.w_fwd[(l * LAYER_ELEMS) / 8 + (r * 2 + 1) * 64 + lane], b[1], z);
            }
            uint32_t mask = 0;
#pragma unroll
            for(int r = 0; r < 4; ++r)
#pragma unroll
                for(int i = 0; i < 4; ++i)
                {
                    const bool pos = acc[r][i] > 0.0f;
                    mask |= (pos ? 1u : 0u) << (4 * r + i);
                    acc[r][i] = pos ? acc[r][i] : 0.0f;
                    act[l + 1][16 * r + 4 * group + i][m_local] = static_cast<__bf16>(acc[r][i]);
                }
            relu_mask[l] = mask;
            pack_next(acc, b);
        }
        f32x4 y = {0.f, 0.f, 0.f, 0.f};
        y = mfma(c.w_fwd[(N_HIDDEN * LAYER_ELEMS) / 8 + 0 * 64 + lane], b[0], y);
        y = mfma(c.w_fwd[(N_HIDDEN * LAYER_ELEMS) / 8 + 1 * 64 + lane], b[1], y);

        // ---- relative L2 loss over rgb; lane group 0 holds outputs 0..3 ----------------------------
        float dy[4] = {0.f, 0.f, 0.f, 0.f};
        if(group == 0 && valid)
        {
            const float4 tg = targets[m];
            const float ybar = ((y[0] + y[1]) + y[2]) / 3.0f;
            const float denom = ybar * ybar + LOSS_EPS;
            const float d0 = y[0] - tg.x, d1 = y[1] - tg.y, d2 = y[2] - tg.z;
            loss_lane += ((d0 * d0) / denom + (d1 * d1) / denom) + (d2 * d2) / denom;
            const float nf = static_cast<float>(n);
            dy[0] = (2.0f * d0 / denom) / nf;
            dy[1] = (2.0f * d1 / denom) / nf;
            dy[2] = (2.0f * d2 / denom) / nf;
        }
#pragma unroll
        for(int i = 0; i < 4; ++i) dzs[4 * group + i][m_local] = static_cast<__bf16>(dy[i]);
        bf16x8 bdy;
#pragma unroll
        for(int j = 0; j < 8; ++j) bdy[j] = static_cast<__bf16>(j < 4 ? dy[j] : 0.0f);
        __syncthreads();

        // ---- W6 gradient: wave w owns input column tile w ------------------------------------------
#pragma unroll
        for(int s = 0; s < 2; ++s)
        {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&dzs[col][32 * s + 8 * group]);
            const bf16x8 bb = *reinterpret_cast<const bf16x8 *>(&act[N_HIDDEN][16 * wave + col][32 * s + 8 * group]);
            dw6 = mfma(a, bb, dw6);
        }
        // dh5 = W6^T dy
        f32x4 dh[4];
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            f32x4 z = {0.f, 0.f, 0.f, 0.f};
            dh[r] = mfma(c.w_bwd[(N_HIDDEN * LAYER_ELEMS) / 8 + r * 64 + lane], bdy, z);
        }

        // ---- hidden layers, W5 .. W1 ---------------------------------------------------------------
#pragma unroll
        for(int li = (int)N_HIDDEN - 1; li >= 0; --li)
        {
            // dz of layer li + 1's input: mask with the ReLU of h_{li + 1}
            const uint32_t mask = relu_mask[li];
#pragma unroll
            for(int r = 0; r < 4; ++r)
#pragma unroll
                for(int i = 0; i < 4; ++i) dh[r][i] = ((mask >> (4 * r + i)) & 1u) ? dh[r][i] : 0.0f;
            __syncthreads();// readers of dzs (previous gradient) are done
#pragma unroll
            for(int r = 0; r < 4; ++r)
#pragma unroll
                for(int i = 0; i < 4; ++i) dzs[16 * r + 4 * group + i][m_local] = static_cast<__bf16>(dh[r][i]);
            pack_next(dh, b);
            __syncthreads();
            // weight gradient of W_{li + 1}: rows = outputs (tile = wave), columns = inputs a_li
#pragma unroll
            for(int q = 0; q < 4; ++q)
#pragma unroll
                for(int s = 0; s < 2; ++s)
                {
                    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&dzs[16 * wave + col][32 * s + 8 * group]);
                    const bf16x8 bb = *reinterpret_cast<const bf16x8 *>(&act[li][16 * q + col][32 * s + 8 * group]);
                    dw[li][q] = mfma(a, bb, dw[li][q]);
                }
            // input gradient: W_{li + 1}^T dz
#pragma unroll
            for(int r = 0; r < 4; ++r)
            {
                f32x4 z = {0.f, 0.f, 0.f, 0.f};
                z = mfma(c.w_bwd[(li * LAYER_ELEMS) / 8 + (r * 2 + 0) * 64 + lane], b[0], z);
                dh[r] = mfma(c.w_bwd[(li * LAYER_ELEMS) / 8 + (r * 2 + 1) * 64 + lane], b[1], z);
            }
        }

        // ---- hash-grid gradient: dh now holds dL/dx0; lane groups 0 and 1 carry the grid features -----
        if(valid && group < 2)
        {
            float p[3];
            unit_position(c.g, rec[0], p);
#pragma unroll
            for(int r = 0; r < 4; ++r)
#pragma unroll
                for(int h = 0; h < 2; ++h)
                {
                    const uint32_t level = 4 * r + 2 * group + h;
                    const float g0 = dh[r][2 * h], g1 = dh[r][2 * h + 1];
                    if(g0 == 0.0f && g1 == 0.0f) continue;
                    uint32_t idx[8];
                    float w[8];
                    level_corners(c.g, level, p, idx, w);
                    const float s0 = g0 * grad_scale, s1 = g1 * grad_scale;
#pragma unroll
                    for(int k = 0; k < 8; ++k)
                    {
                        const half2v val = {static_cast<_Float16>(w[k] * s0), static_cast<_Float16>(w[k] * s1)};
                        __builtin_amdgcn_global_atomic_fadd_v2f16(grid_grad + (level * table + idx[k]), val);
                    }
                }
        }
    }

    // ---- per-block weight gradients, logical [out][in] layout ------------------------------------------
    float *part = partials + (size_t)blockIdx.x * N_MLP;
#pragma unroll
    for(uint32_t li = 0; li < N_HIDDEN; ++li)
#pragma unroll
        for(int q = 0; q < 4; ++q)
#pragma unroll
            for(int i = 0; i < 4; ++i)
            {
                const uint32_t out = 16 * wave + 4 * group + i, in = 16 * q + col;
                part[layer_offset(li) + out * WIDTH + in] = dw[li][q][i];
            }
#pragma unroll
    for(int i = 0; i < 4; ++i)
    {
        const uint32_t out = 4 * group + i, in = 16 * wave + col;
        part[layer_offset(N_HIDDEN) + out * WIDTH + in] = dw6[i];
    }
    // loss: one atomic per wave
    for(int o = 32; o > 0; o >>= 1) loss_lane += __shfl_xor(loss_lane, o);
    if(lane == 0) atomicAdd(loss_sum, loss_lane);
}

struct adam_consts
{
    float lr, beta1, beta2, eps, c1, c2;// c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t)
};

__device__ __forceinline__ void adam_update(float &p, float &m, float &v, float g, const adam_consts &a)
{
    m = a.beta1 * m + (1.0f - a.beta1) * g;
    v = a.beta2 * v + (1.0f - a.beta2) * (g * g);
    p = p - (a.lr * (m * a.c1)) / (sqrtf(v * a.c2) + a.eps);
}

//! write W's bf16 copies (forward and transposed swizzles) for weight k of the flat master array
__device__ __forceinline__ void publish_weight(uint32_t k, float w, __bf16 *w_fwd, __bf16 *w_bwd)
{
    const uint32_t layer = k / LAYER_ELEMS, rem = k % LAYER_ELEMS;
    const uint32_t out = rem / WIDTH, in = rem % WIDTH;
    w_fwd[fwd_pos(layer, out, in)] = static_cast<__bf16>(w);
    w_bwd[bwd_pos(layer, out, in)] = static_cast<__bf16>(w);
}

__global__ void k_rc_adam_mlp(float *__restrict__ master, float *__restrict__ mom, float *__restrict__ vel,
                              float *__restrict__ grad_out, const float *__restrict__ partials, uint32_t n_blocks,
                              adam_consts a, int apply, __bf16 *__restrict__ w_fwd, __bf16 *__restrict__ w_bwd)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= N_MLP) return;
    // eight independent partial sums (blocks b = j mod 8), combined in a fixed order: deterministic
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t b = 0;
    for(; b + 8 <= n_blocks; b += 8)
#pragma unroll
        for(int j = 0; j < 8; ++j) acc[j] = acc[j] + partials[(size_t)(b + j) * N_MLP + k];
#pragma unroll
    for(int j = 0; j < 8; ++j)
        if(b + j < n_blocks) acc[j] = acc[j] + partials[(size_t)(b + j) * N_MLP + k];
    const float g = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    grad_out[k] = g;
    if(!apply || g == 0.0f) return;
    float p = master[k], m = mom[k], v = vel[k];
    adam_update(p, m, v, g, a);
    master[k] = p;
    mom[k] = m;
    vel[k] = v;
    publish_weight(k, p, w_fwd, w_bwd);
}

__global__ void k_rc_publish_mlp(const float *__restrict__ master, __bf16 *__restrict__ w_fwd, __bf16 *__restrict__ w_bwd)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k < N_MLP) publish_weight(k, master[k], w_fwd, w_bwd);
}

//! (the gradient entries it reads as non-zero are set back to zero, so the next step starts from a clear buffer
//! without a fill)
__global__ void k_rc_adam_grid(float *__restrict__ master, float *__restrict__ mom, float *__restrict__ vel,
                               _Float16 *__restrict__ grad, float inv_scale, __half *__restrict__ stored,
                               uint64_t n, adam_consts a)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n) return;
    const _Float16 raw = grad[k];
    if(raw == static_cast<_Float16>(0.0f)) return;
    grad[k] = static_cast<_Float16>(0.0f);
    const float g = static_cast<float>(raw) * inv_scale;
    // untouched entries keep their parameters and moments; a non-finite (overflowed) gradient is dropped
    if(g == 0.0f || !isfinite(g)) return;
    float p = master[k], m = mom[k], v = vel[k];
    adam_update(p, m, v, g, a);
    master[k] = p;
    mom[k] = m;
    vel[k] = v;
    stored[k] = __float2half_rn(p);
}

__global__ void k_rc_store_grid(const float *__restrict__ master, __half *__restrict__ stored, uint64_t n)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(k < n) stored[k] = __float2half_rn(master[k]);
}

//! uniform U(-1e-4, 1e-4) from a SplitMix64 hash of (seed, index)
__global__ void k_rc_init_grid(float *__restrict__ master, uint64_t n, uint32_t seed)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n) return;
    uint64_t z = ((uint64_t)seed << 40) + k + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const float u = static_cast<float>(z >> 40) * (1.0f / 16777216.0f);
    master[k] = (2.0f * u - 1.0f) * 1e-4f;
}

__global__ void k_rc_fill(float *__restrict__ p, uint64_t n, float value)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(k < n) p[k] = value;
}

static void fill(hipStream_t s, float *p, uint64_t n, float value)
{
    if(n) k_rc_fill<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(p, n, value);
}

static uint32_t train_blocks(vkrt_rc *c, uint32_t n)
{
    // one 4-wave block per CU (the kernel runs one wave per SIMD)
    const uint32_t iterations = div_up(n, TRAIN_SAMPLES);
    return std::max<uint32_t>(1u, std::min<uint32_t>(iterations, (uint32_t)c->ctx->cu_count));
}

//! power-of-two loss scale of the fp16 grid gradient: per-sample gradients scale as 1 / n, so the scaled
//! contributions stay near 2^-6 (inside fp16's normal range, far from its 65504 limit)
float grid_grad_scale(uint32_t n)
{
    const int e = n ? 31 - __builtin_clz(n) : 0;
    return std::ldexp(1.0f, std::min(12, std::max(-6, e - 6)));
}

__global__ void k_rc_grid_grad_f32(const _Float16 *__restrict__ g, float inv_scale, float *__restrict__ out, uint64_t n)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(k < n) out[k] = static_cast<float>(g[k]) * inv_scale;
}

void train(vkrt_rc *c, const float4 *records, uint32_t stride_f4, const float4 *targets, uint32_t n, bool grad_only,
           const uint32_t *n_dev, hipStream_t s)
{
    const uint32_t nb = train_blocks(c, n);
    c->partials.reserve((size_t)nb * N_MLP * 4);
    c->grad_scale = grid_grad_scale(n);
    if(!c->grid_grad_clear) fill(s, c->grid_grad.as<float>(), c->n_grid / 2, 0.0f);// n_grid fp16 values
    // a gradient-only step leaves its gradients for reading; a full step's Adam pass clears them
    c->grid_grad_clear = !grad_only;
    fill(s, c->loss.as<float>(), 1, 0.0f);
    if(n)
        k_rc_train<<<nb, TRAIN_BLOCK, 0, s>>>(device_view(c), records, stride_f4, targets, n, n_dev,
                                              c->grid_grad.as<half2v>(), c->grad_scale, c->partials.as<float>(),
                                              c->loss.as<float>());
    else
        fill(s, c->partials.as<float>(), (uint64_t)nb * N_MLP, 0.0f);
    adam_consts a{};
    if(!grad_only)
    {
        c->step += 1;
        const vkrt_rc_params &p = c->params;
        a = adam_consts{p.learning_rate, p.beta1, p.beta2, p.epsilon,
                        (float)(1.0 / (1.0 - std::pow((double)p.beta1, (double)c->step))),
                        (float)(1.0 / (1.0 - std::pow((double)p.beta2, (double)c->step)))};
    }
    k_rc_adam_mlp<<<div_up(N_MLP, 256), 256, 0, s>>>(c->mlp_master.as<float>(), c->mlp_m.as<float>(),
                                                       c->mlp_v.as<float>(), c->mlp_grad.as<float>(),
                                                       c->partials.as<float>(), n ? nb : 1, a, grad_only ? 0 : 1,
                                                       c->w_fwd.as<__bf16>(), c->w_bwd.as<__bf16>());
    if(!grad_only)
        k_rc_adam_grid<<<(uint32_t)((c->n_grid + 255) / 256), 256, 0, s>>>(
            c->grid_master.as<float>(), c->grid_m.as<float>(), c->grid_v.as<float>(), c->grid_grad.as<_Float16>(),
            1.0f / c->grad_scale, c->grid_half.as<__half>(), c->n_grid, a);
    VKRT_HIP(hipGetLastError());
}

void query(vkrt_rc *c, const float4 *records, uint32_t stride_f4, uint32_t n, float4 *out, float4 *pixel_acc,
           const uint32_t *count_on_device)
{
    if(n == 0) return;
    const uint32_t per_block = QUERY_BLOCK / 64 * 16 * QUERY_TILES;
    k_rc_query<<<div_up(n, per_block), QUERY_BLOCK, 0, c->ctx->stream>>>(device_view(c), records, stride_f4, n,
                                                                         count_on_device, out, pixel_acc);
    VKRT_HIP(hipGetLastError());
}

static void reset_optimizer(vkrt_rc *c)
{
    hipStream_t s = c->ctx->stream;
    fill(s, c->grid_m.as<float>(), c->n_grid, 0.0f);
    fill(s, c->grid_v.as<float>(), c->n_grid, 0.0f);
    fill(s, c->mlp_m.as<float>(), N_MLP, 0.0f);
    fill(s, c->mlp_v.as<float>(), N_MLP, 0.0f);
    c->step = 0;
}

static void publish(vkrt_rc *c)
{
    hipStream_t s = c->ctx->stream;
    k_rc_store_grid<<<(uint32_t)((c->n_grid + 255) / 256), 256, 0, s>>>(c->grid_master.as<float>(),
                                                                         c->grid_half.as<__half>(), c->n_grid);
    k_rc_publish_mlp<<<div_up(N_MLP, 256), 256, 0, s>>>(c->mlp_master.as<float>(), c->w_fwd.as<__bf16>(),
                                                          c->w_bwd.as<__bf16>());
    VKRT_HIP(hipGetLastError());
}

//! Glorot-uniform MLP weights from SplitMix64(seed); W6 rows 3..15 (unused outputs) are zero
static std::vector<float> init_mlp(uint32_t seed)
{
    std::vector<float> w(N_MLP, 0.0f);
    uint64_t x = 0x5DEECE66Dull ^ ((uint64_t)seed << 17);
    auto next = [&x]() {
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        return static_cast<double>(z >> 40) / 16777216.0;
    };
    for(uint32_t l = 0; l < N_LAYERS; ++l)
    {
        const uint32_t rows = l < N_HIDDEN ? WIDTH : OUTPUTS;
        const double bound = std::sqrt(6.0 / (double)(WIDTH + (l < N_HIDDEN ? WIDTH : 3)));
        for(uint32_t o = 0; o < rows; ++o)
            for(uint32_t i = 0; i < WIDTH; ++i)
            {
                const float v = (float)((2.0 * next() - 1.0) * bound);
                if(l < N_HIDDEN || o < 3) w[layer_offset(l) + o * WIDTH + i] = v;
            }
    }
    return w;
}

static grid_desc make_desc(const vkrt_rc_params &p)
{
    grid_desc d{};
    const double b = std::exp((std::log((double)p.max_resolution) - std::log((double)p.base_resolution)) / (LEVELS - 1));
    const uint64_t t = uint64_t(1) << p.log2_table;
    for(uint32_t l = 0; l < LEVELS; ++l)
    {
        d.res[l] = (uint32_t)std::floor((double)p.base_resolution * std::pow(b, (double)l));
        const uint64_t r1 = d.res[l] + 1ull;
        if(r1 * r1 * r1 <= t) d.dense_mask |= 1u << l;
    }
    d.log2_t = p.log2_table;
    for(int a = 0; a < 3; ++a)
    {
        d.bmin[a] = 0.0f;
        d.inv_extent[a] = 1.0f;
    }
    return d;
}

}// namespace vkrt::rc

using namespace vkrt;
using namespace vkrt::rc;

extern "C" {

void vkrt_rc_default_params(vkrt_rc_params *out)
{
    if(!out) return;
    *out = {};
    out->log2_table = 19;
    out->base_resolution = 16;
    out->max_resolution = 2048;
    out->learning_rate = 1e-2f;
    out->beta1 = 0.9f;
    out->beta2 = 0.99f;
    out->epsilon = 1e-15f;
}

int vkrt_rc_create(vkrt_ctx *ctx, const vkrt_rc_params *params, uint32_t seed, vkrt_rc **out)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(out, "null output handle");
        *out = nullptr;
        auto c = std::make_unique<vkrt_rc>();
        c->ctx = ctx;
        if(params) c->params = *params;
        else vkrt_rc_default_params(&c->params);
        const vkrt_rc_params &p = c->params;
        VKRT_REQUIRE(p.log2_table >= 10 && p.log2_table <= 24, "log2_table must be in 10..24");
        VKRT_REQUIRE(p.base_resolution >= 2 && p.max_resolution >= p.base_resolution && p.max_resolution <= (1u << 20),
                     "resolutions must satisfy 2 <= base <= max <= 2^20");
        VKRT_REQUIRE(p.learning_rate > 0.f && p.beta1 >= 0.f && p.beta1 < 1.f && p.beta2 >= 0.f && p.beta2 < 1.f &&
                         p.epsilon >= 0.f,
                     "Adam parameters out of range");
        c->desc = make_desc(p);
        c->n_grid = (uint64_t)LEVELS * (1ull << p.log2_table) * 2ull;
        c->grid_master.reserve(c->n_grid * 4);
        c->grid_half.reserve(c->n_grid * 2);
        c->grid_grad.reserve(c->n_grid * 2);
        c->grid_grad_f32.reserve(c->n_grid * 4);
        c->grid_m.reserve(c->n_grid * 4);
        c->grid_v.reserve(c->n_grid * 4);
        c->mlp_master.reserve(N_MLP * 4);
        c->mlp_grad.reserve(N_MLP * 4);
        c->mlp_m.reserve(N_MLP * 4);
        c->mlp_v.reserve(N_MLP * 4);
        c->w_fwd.reserve(N_LAYERS * LAYER_ELEMS * 2);
        c->w_bwd.reserve(N_LAYERS * LAYER_ELEMS * 2);
        c->loss.reserve(256);
        hipStream_t s = ctx->stream;
        // padding slots of the transposed W6 swizzle stay zero
        VKRT_HIP(hipMemsetAsync(c->w_bwd.ptr, 0, N_LAYERS * LAYER_ELEMS * 2, s));
        VKRT_HIP(hipMemsetAsync(c->w_fwd.ptr, 0, N_LAYERS * LAYER_ELEMS * 2, s));
        fill(s, c->mlp_grad.as<float>(), N_MLP, 0.0f);
        k_rc_init_grid<<<(uint32_t)((c->n_grid + 255) / 256), 256, 0, s>>>(c->grid_master.as<float>(), c->n_grid, seed);
        const std::vector<float> w = init_mlp(seed);
        VKRT_HIP(hipMemcpyAsync(c->mlp_master.ptr, w.data(), N_MLP * 4, hipMemcpyHostToDevice, s));
        reset_optimizer(c.get());
        publish(c.get());
        VKRT_HIP(hipStreamSynchronize(s));
        *out = c.release();
    });
}

void vkrt_rc_destroy(vkrt_rc *c)
{
    if(!c) return;
    (void)hipSetDevice(c->ctx->device);
    (void)hipStreamSynchronize(c->ctx->stream);
    delete c;
}

int vkrt_rc_get_config(const vkrt_rc *c, vkrt_rc_params *params, uint32_t *n_mlp, uint64_t *n_grid)
{
    if(!c) return VKRT_E_INVALID_ARG;
    if(params) *params = c->params;
    if(n_mlp) *n_mlp = N_MLP;
    if(n_grid) *n_grid = c->n_grid;
    return VKRT_OK;
}

int vkrt_rc_set_bounds(vkrt_rc *c, const float *bmin, const float *bmax)
{
    if(!c) return VKRT_E_INVALID_ARG;
    return guarded(c->ctx, [&] {
        VKRT_REQUIRE(bmin && bmax, "null bounds");
        for(int a = 0; a < 3; ++a)
        {
            VKRT_REQUIRE(std::isfinite(bmin[a]) && std::isfinite(bmax[a]) && bmax[a] > bmin[a],
                         "bounds must be finite with max > min");
            c->desc.bmin[a] = bmin[a];
            c->desc.inv_extent[a] = 1.0f / (bmax[a] - bmin[a]);
        }
    });
}

int vkrt_rc_query(vkrt_rc *c, const float *records, uint32_t n, float *out, int flags)
{
    if(!c) return VKRT_E_INVALID_ARG;
    vkrt_ctx *ctx = c->ctx;
    return guarded(ctx, [&] {
        VKRT_REQUIRE(n == 0 || (records && out), "null records or output");
        const float4 *rec = reinterpret_cast<const float4 *>(input(ctx, 0, records, (size_t)n * 16, flags));
        float4 *o = reinterpret_cast<float4 *>(output(ctx, 1, out, (size_t)n * 4, flags));
        rc::query(c, rec, RECORD_F4, n, o, nullptr, nullptr);
        finish_output(ctx, 1, out, (size_t)n * 4, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_rc_train(vkrt_rc *c, const float *records, const float *targets, uint32_t n, int flags, float *loss)
{
    if(!c) return VKRT_E_INVALID_ARG;
    vkrt_ctx *ctx = c->ctx;
    return guarded(ctx, [&] {
        VKRT_REQUIRE(n == 0 || (records && targets), "null records or targets");
        const float4 *rec = reinterpret_cast<const float4 *>(input(ctx, 0, records, (size_t)n * 16, flags));
        const float4 *tg = reinterpret_cast<const float4 *>(input(ctx, 1, targets, (size_t)n * 4, flags));
        rc::train(c, rec, RECORD_F4, tg, n, (flags & VKRT_RC_GRAD_ONLY) != 0, nullptr, c->ctx->stream);
        if(loss)
        {
            float sum = 0.f;
            VKRT_HIP(hipMemcpyAsync(&sum, c->loss.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
            VKRT_HIP(hipStreamSynchronize(ctx->stream));
            *loss = n ? sum / (float)n : 0.f;
        }
        finish_call(ctx, flags);
    });
}

int vkrt_rc_get_params(vkrt_rc *c, float *mlp, float *grid)
{
    if(!c) return VKRT_E_INVALID_ARG;
    return guarded(c->ctx, [&] {
        hipStream_t s = c->ctx->stream;
        if(mlp) VKRT_HIP(hipMemcpyAsync(mlp, c->mlp_master.ptr, N_MLP * 4, hipMemcpyDeviceToHost, s));
        if(grid) VKRT_HIP(hipMemcpyAsync(grid, c->grid_master.ptr, c->n_grid * 4, hipMemcpyDeviceToHost, s));
        VKRT_HIP(hipStreamSynchronize(s));
    });
}

int vkrt_rc_set_params(vkrt_rc *c, const float *mlp, const float *grid)
{
    if(!c) return VKRT_E_INVALID_ARG;
    return guarded(c->ctx, [&] {
        hipStream_t s = c->ctx->stream;
        if(mlp)
        {
            std::vector<float> w(mlp, mlp + N_MLP);
            for(uint32_t o = 3; o < OUTPUTS; ++o)
                for(uint32_t i = 0; i < WIDTH; ++i) w[layer_offset(N_HIDDEN) + o * WIDTH + i] = 0.0f;
            VKRT_HIP(hipMemcpyAsync(c->mlp_master.ptr, w.data(), N_MLP * 4, hipMemcpyHostToDevice, s));
            VKRT_HIP(hipStreamSynchronize(s));
        }
        if(grid) VKRT_HIP(hipMemcpyAsync(c->grid_master.ptr, grid, c->n_grid * 4, hipMemcpyHostToDevice, s));
        reset_optimizer(c);
        publish(c);
        VKRT_HIP(hipStreamSynchronize(s));
    });
}

int vkrt_rc_get_gradients(vkrt_rc *c, float *mlp, float *grid)
{
    if(!c) return VKRT_E_INVALID_ARG;
    return guarded(c->ctx, [&] {
        hipStream_t s = c->ctx->stream;
        if(mlp) VKRT_HIP(hipMemcpyAsync(mlp, c->mlp_grad.ptr, N_MLP * 4, hipMemcpyDeviceToHost, s));
        if(grid)
        {
            k_rc_grid_grad_f32<<<(uint32_t)((c->n_grid + 255) / 256), 256, 0, s>>>(
                c->grid_grad.as<_Float16>(), 1.0f / c->grad_scale, c->grid_grad_f32.as<float>(), c->n_grid);
            VKRT_HIP(hipGetLastError());
            VKRT_HIP(hipMemcpyAsync(grid, c->grid_grad_f32.ptr, c->n_grid * 4, hipMemcpyDeviceToHost, s));
        }
        VKRT_HIP(hipStreamSynchronize(s));
    });
}

uint32_t vkrt_rc_step(const vkrt_rc *c) { return c ? c->step : 0u; }

}// extern "C"
