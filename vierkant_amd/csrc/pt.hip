// pt.hip - wavefront path tracer with PBRPathTracer's settings and accumulation semantics.
//
// Reference: the raygen/closest-hit/miss programs of shaders/slang/raypipeline.slang (:91-270, :360-680,
// :746-761) run by the Vulkan RT pipeline, and the host state machine of src/PBRPathTracer.cpp
// (render_scene :184-255, reset_accumulator :677-680, per-frame seed :588). On gfx950 the per-pixel
// sample x bounce loop becomes a wavefront over all pixels of one sample:
//
//   k_pt_begin   per pixel: first sample draws the pixel's RNG state and sample offset; every sample draws
//                the lens sample and writes the camera ray into the SoA ray queue
//   k_pt_trace   per queued path: closest hit through the BVH, back faces culled while the path is in air
//   k_pt_shade   per queued path: the bounce's RNG hash, miss (sky) or closest-hit shading (emission,
//                sample_disney, throughput, media stack, Russian roulette); survivors are appended to the
//                next queue with one wave ballot + one atomic per wave; finished paths add their radiance
//                to the pixel's sample sum
//   k_pt_resolve per pixel: mean over the batch's samples, running mean over batches, pixel_data_t
//
// All per-path state is SoA in HBM (float4 / uint4 per path) so every kernel reads and writes it
// coalesced; queues hold pixel indices.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "bvh_types.hpp"
#include "context.hpp"
#include "radiance_cache.hpp"
#include "scene_walk.hpp"
#include "trace.hpp"
#include "traverse.hpp"
#include "vierkant/pt_math.hpp"

using namespace vierkant::pt;
using vierkant::rt::vec3;

namespace vkrt
{

constexpr uint32_t MEDIA_STACK = 4;          // MAX_MEDIA_STACK_SIZE (raypipeline.slang:21)
constexpr uint32_t NO_OBJECT = 0xFFFFu;      // utils::MAX_UINT16
constexpr uint32_t F_MEDIA_MASK = 3u, F_REGULARIZE = 4u, F_CAUSTIC = 8u, F_FIRST_MISS = 16u, F_TRAIN = 32u;
constexpr uint32_t QUERY_F4 = 5;   // cache query record: 4 float4 of features + (pixel, beta)
constexpr uint32_t VERTEX_F4 = 7;  // training vertex: 4 float4 of features + (emission, kind) + (weight, -)
                                   // + (light bodies met on the segment arriving here, -)
constexpr uint32_t SHADOW_F4 = 6;  // shadow record: ray o, d; (radiance, mis weight); (F, cos); (beta, pdf);
                                   // (pixel, training vertex, -, -)
constexpr uint32_t NO_VERTEX = 0xffffffffu;
constexpr uint32_t TRAIN_COUNTERS = 8;// offset of the training stream's counters
constexpr int PT_BLOCK = 256;
constexpr int PT_LDS = 16;
// the closest-hit trace of the queued paths: the trace kernels' launch shape (64 lanes per workgroup; the
// workgroups one XCD receives take PT_TRACE_CHUNK consecutive queue blocks in turn)
constexpr int PT_TRACE_BLOCK = 64;
constexpr uint32_t PT_TRACE_CHUNK = 8;
// Per-depth launches cover every pixel (a workgroup per tile, those past the live queue return at once). The
// persistent alternative (VKRT_PT_LAUNCH_PERSISTENT) launches the workgroups the chip holds at once, the trace
// kernel claiming tiles of PT_TRACE_BLOCK paths from per-XCD queues at counter + PT_CLAIMS and the shadow
// kernel striding over its records; measured on the room scene (1920x1080, 64 spp, depth 8) it is slower,
// 230 ms against 192 ms a frame: the hardware dispatcher hands out tiles faster than the claim atomics.
constexpr uint32_t PT_CLAIMS = 64;            // word offset of the trace kernel's claim counters
constexpr uint32_t PT_COUNTER_WORDS = PT_CLAIMS + CLAIM_COUNTERS;
constexpr uint32_t PT_TRACE_WAVES_PER_SIMD = 6;// the trace kernel's register cap (amdgpu_waves_per_eu): 80 VGPRs
// the filter-free trace kernel fits 64 VGPRs with no spill: 8 waves per SIMD instead of 7 (room, 64 spp at
// 1080p: 170.9 ms against 174.2)
constexpr uint32_t PT_TRACE_PLAIN_WAVES_PER_SIMD = 8;
// the filter-free flat shadow kernel at 5 waves per SIMD: 96 VGPRs and 21 spilled instead of 121 at 4 (room
// with six lights, 16 spp at 1080p: 90.7 ms against 93.4; at 6 waves 70 registers spill)
constexpr uint32_t PT_SHADOW_PLAIN_WAVES_PER_SIMD = 5;
// the two-level walk's cap: 96 VGPRs and 5 spilled instead of 80 and 14 (instanced room, 16 spp at 1080p:
// 68.2 ms against 69.9 at 6 waves and 70.5 at 4)
constexpr uint32_t PT_TRACE_INST_WAVES_PER_SIMD = 5;
constexpr uint32_t PT_TAIL_DEPTH = VKRT_PT_TAIL_DEPTH_DEFAULT;// survivors at this depth finish in one lane each

//! f32tof16 / f16tof32 round trip (round to nearest even) of the media stack's packed ior
__host__ __device__ inline float round_half(float x)
{
    uint32_t b;
    __builtin_memcpy(&b, &x, 4);
    const uint32_t sign = b & 0x80000000u;
    const uint32_t absb = b & 0x7fffffffu;
    if(absb >= 0x47800000u) return x;// |x| >= 65536: out of the ior range the tracer uses
    if(absb < 0x38800000u)          // below the smallest normal half: keep as is (never an ior)
        return x;
    // keep 10 mantissa bits, round to nearest even
    uint32_t lsb = (absb >> 13) & 1u;
    uint32_t r = absb + 0xfffu + lsb;
    r &= ~0x1fffu;
    r |= sign;
    float out;
    __builtin_memcpy(&out, &r, 4);
    return out;
}

//! f32tof16 / f16tof32: the value as binary16 stores it (round to nearest even, overflow to infinity)
__device__ inline float half_value(float x) { return static_cast<float>(static_cast<_Float16>(x)); }

struct gpu_material
{
    float4 color;    // rgb, alpha
    float4 emission; // rgb, strength
    float4 mrti;     // metalness, roughness, transmission, ior
    float4 cc;       // clearcoat, clearcoat_roughness, sheen_roughness, diffuse_transmission
    float4 sheen;    // sheen color rgb, two_sided flag
    float4 dtc;      // diffuse transmission color rgb, -
    uint4 tex;       // 1 + color texture (0: none), alpha mode, alpha cutoff bits, -
    uint4 tex2;      // 1 + normal, emission, ao/rough/metal, transmission texture
    uint4 tex3;      // 1 + diffuse transmission, diffuse transmission color texture, -, -
    float4 vol;      // interior medium: sigma_t rgb = -log(attenuation color) / attenuation distance, 1 if any
};

//! a transmissive candidate's contribution to a shadow ray's Beer-Lambert optical depth
//! (visibility_test.slang:123-129): -sigma_t t where the ray enters the material's volume (front face), +sigma_t t
//! where it leaves; km carries the front-face flag in its top bit
constexpr uint32_t KM_FRONT = 0x80000000u;
__device__ __forceinline__ uint32_t candidate_tag(uint32_t mid, const gpu_material &m, const vec3 &d, const float4 &p0,
                                                  const float4 &p1, const float4 &p2)
{
    if(!(m.vol.w > 0.0f)) return mid;
    const vec3 e1{p0.w, p1.x, p1.y}, e2{p1.z, p1.w, p2.x};
    return dot(e1, vierkant::rt::cross(d, e2)) > 0.0f ? (mid | KM_FRONT) : mid;
}

constexpr uint32_t MAX_TEX_LEVELS = 16;

//! one texture's mip chain inside the shared texel array (RGBA8 packed in a uint32)
struct tex_desc
{
    uint32_t width, height, levels;
    uint32_t sampler;// bits 0-1 address mode u, 2-3 address mode v, 4 min filter nearest, 5 mag filter nearest
    uint32_t offset[MAX_TEX_LEVELS];
};

struct pt_frame
{
    uint32_t width, height, n_pix;
    uint32_t num_samples, max_depth;
    uint32_t seed, batch_index, disable_material;
    float max_path_beta, environment;
    // radiance cache: paths that are not training paths stop at hit depth cache_depth and query the cache;
    // every train_stride-th pixel (from train_offset) traces a full training path
    uint32_t cache_on, cache_depth, train_stride, train_offset, n_train_paths, train_max;
    // analytic lights: next-event estimation and the light bodies BSDF rays meet
    uint32_t n_lights, mis_mode, suppress_caustics;
    uint32_t any_volume;// some material has an interior medium: paths keep the media stack's materials
    float cone_spread;// atan(2 tan(fov / 2) / height) as a half (raypipeline.slang:102)
    vec3 eye, right, up, forward;
    float tan_half_x, tan_half_y, aperture, focal_distance;
};

struct pt_buffers
{
    float4 *ray_o, *ray_d;  // origin + tmin, direction + tmax
    uint4 *hit;             // t, u, v, prim | backface << 31
    float4 *beta, *rad;     // throughput, path radiance
    uint4 *state;           // raygen rng, flags, sample offset bits, -
    float4 *iors;           // media stack iors
    uint2 *media;           // media stack: the material of each level's interior, 16 bits each (0xffff: none)
    float4 *acc;            // sum of the batch's sample radiance (rgb), coverage
    uint4 *info;            // first-hit depth bits, normal, object id, albedo
    uint32_t *queue[2];
    uint32_t *counter;      // [0], [1]: queue sizes; [2..3]: segment count (64-bit); [4]: cache queries;
                            // [5]: training records; [6]: shadow rays. Training paths of the cache mode run on
                            // their own stream with the same layout at counter + TRAIN_COUNTERS
    uint32_t *train_queue;  // the training paths' first queue (cache mode)
    vkrt_pixel *pixels;
    float4 *qrec;           // cache queries (QUERY_F4 float4 each)
    float4 *tvtx;           // training paths' vertices [path][depth] (VERTEX_F4 float4 each)
    float4 *trec, *ttgt;    // training records (4 float4) and targets
    float4 *shadow;         // next-event shadow rays of one depth (SHADOW_F4 float4 each), count in counter[6]
    float2 *cone;           // ray cone per path: spread angle, width
    uint32_t *hit_inst;     // instanced scenes: the instance of each path's hit
};

struct pt_scene_dev
{
    const float *pos, *nrm;
    const uint32_t *idx, *tri_object, *object_material;
    const float *object_tmax;
    const gpu_material *mats;
    const vkrt_light *lights;
    const float2 *uv;         // texture coordinates per vertex
    const tex_desc *tex;      // texture descriptors
    const uint32_t *texels;   // RGBA8 mip levels of all textures
    const float *tan;         // tangents per vertex (normal maps)
    // instanced scenes: per instance the shading transform (transform_t: rotation quaternion, translation and
    // x scale, y and z scale) and the world-to-object rows; per object its object-space box half extents
    const float4 *inst_xf;
    const inst_dev *insts;
    const float *object_he;
};

//! the instanced scene's top level for the trace kernels (unused by the flat-scene instances)
struct pt_tlas
{
    const vkrt_bvh4_node *tlas;
    int32_t root;
    uint32_t empty;
    const uint32_t *order;
    const inst_dev *insts;
    const blas_dev *blases;
};

//! transform_t applied to a point and to a direction (utils/transform.slang: rotate_quat, apply_transform)
__device__ __forceinline__ vec3 rotate_quat(const float4 &q, const vec3 &v)
{
    const vec3 qv{q.x, q.y, q.z};
    const vec3 inner = vierkant::rt::cross(qv, v) + v * q.w;
    return v + vierkant::rt::cross(qv, inner) * 2.0f;
}

__device__ __forceinline__ vec3 apply_instance(const float4 *xf, const vec3 &p)
{
    const float4 q = xf[0], ts = xf[1], s2 = xf[2];
    const vec3 scaled{p.x * ts.w, p.y * s2.x, p.z * s2.y};
    const vec3 r = rotate_quat(q, scaled);
    return vec3{r.x + ts.x, r.y + ts.y, r.z + ts.z};
}

__device__ __forceinline__ vec3 ld3f(const float *p, uint32_t i) { return vec3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// ---------------------------------------------------------------------------------------------
// textures: RGBA8 UNORM, repeat addressing, bilinear within a level and linear between levels, the
// sampler of model_loading.cpp:70 and Material.hpp:175-179 (texture_sampler_t defaults)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float4 unpack_unorm8(uint32_t c)
{
    return make_float4(static_cast<float>(c & 0xFFu) / 255.0f, static_cast<float>((c >> 8) & 0xFFu) / 255.0f,
                       static_cast<float>((c >> 16) & 0xFFu) / 255.0f, static_cast<float>(c >> 24) / 255.0f);
}

//! texel index i of a level n texels wide under an address mode (Vulkan's repeat, mirrored repeat and
//! clamp to edge); i is taken modulo 2n first in 64-bit so that far-away tiles wrap exactly
__device__ __forceinline__ int address_texel(long long i, int n, uint32_t mode)
{
    if(mode == VKRT_ADDRESS_CLAMP_TO_EDGE) return static_cast<int>(i < 0 ? 0 : (i >= n ? n - 1 : i));
    if(mode == VKRT_ADDRESS_MIRRORED_REPEAT)
    {
        const long long m = 2ll * n, t = ((i % m) + m) % m;
        return static_cast<int>(t < n ? t : m - 1 - t);
    }
    return static_cast<int>(((i % n) + n) % n);
}

__device__ __forceinline__ float4 lerp4(const float4 &a, const float4 &b, float t)
{
    return make_float4(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z), a.w + t * (b.w - a.w));
}

//! one level: bilinear (texel centres at (i + 0.5) / n) or nearest (texel floor(u n))
__device__ float4 sample_level(const tex_desc &d, const uint32_t *__restrict__ texels, uint32_t level, float u, float v,
                               bool nearest)
{
    const int w = static_cast<int>(max(1u, d.width >> level)), h = static_cast<int>(max(1u, d.height >> level));
    const uint32_t mu = d.sampler & 3u, mv = (d.sampler >> 2) & 3u;
    const uint32_t *t = texels + d.offset[level];
    if(nearest)
    {
        const int x = address_texel(static_cast<long long>(floorf(u * static_cast<float>(w))), w, mu);
        const int y = address_texel(static_cast<long long>(floorf(v * static_cast<float>(h))), h, mv);
        return unpack_unorm8(t[y * w + x]);
    }
    const float fx = u * static_cast<float>(w) - 0.5f, fy = v * static_cast<float>(h) - 0.5f;
    const float x0f = floorf(fx), y0f = floorf(fy);
    const float ax = fx - x0f, ay = fy - y0f;
    const long long xl = static_cast<long long>(x0f), yl = static_cast<long long>(y0f);
    const int x0 = address_texel(xl, w, mu), x1 = address_texel(xl + 1, w, mu);
    const int y0 = address_texel(yl, h, mv), y1 = address_texel(yl + 1, h, mv);
    const float4 c00 = unpack_unorm8(t[y0 * w + x0]), c10 = unpack_unorm8(t[y0 * w + x1]);
    const float4 c01 = unpack_unorm8(t[y1 * w + x0]), c11 = unpack_unorm8(t[y1 * w + x1]);
    return lerp4(lerp4(c00, c10, ax), lerp4(c01, c11, ax), ay);
}

//! SampleLevel(uv, lambda): lambda clamped to the chain (a NaN or negative level is level 0), the min filter
//! where lambda > 0 and the mag filter elsewhere, levels blended linearly
__device__ float4 sample_texture(const tex_desc &d, const uint32_t *__restrict__ texels, float u, float v, float lambda)
{
    const float top = static_cast<float>(d.levels - 1u);
    const bool nearest = lambda > 0.0f ? (d.sampler & 16u) != 0u : (d.sampler & 32u) != 0u;
    if(!(lambda > 0.0f)) lambda = 0.0f;
    if(lambda > top) lambda = top;
    if(!(fabsf(u) < 1e9f) || !(fabsf(v) < 1e9f)) u = v = 0.0f;// no fraction is left there; keeps texel indices in range
    const uint32_t l0 = static_cast<uint32_t>(lambda);
    const float f = lambda - static_cast<float>(l0);
    const float4 a = sample_level(d, texels, l0, u, v, nearest);
    if(f == 0.0f) return a;
    return lerp4(a, sample_level(d, texels, min(l0 + 1u, d.levels - 1u), u, v, nearest), f);
}

//! rnd() of a copy of an RNG state: the draw a stage makes without advancing its caller's stream
__device__ __forceinline__ float rnd_peek(uint32_t state) { return rnd(state); }

//! the any-hit test of MASK and BLEND materials inside traversal: the candidate's color alpha times its
//! texture's alpha at level 0, kept when not below the cutoff (MASK) or a uniform draw (BLEND).
//! Closest-hit rays (raypipeline.slang:686-720) see disable_material's stand-in color (0.8, 1) and draw once
//! per trace from a copy of the path's RNG state (the same `xi` for every candidate); shadow rays
//! (visibility_test.slang:136-155) see the material as it is and draw per candidate, here from
//! xxhash32(seed, triangle) so that the outcome does not depend on the traversal order.
struct alpha_filter
{
    static constexpr bool active = true;
    const uint32_t *idx;
    const float2 *uv;
    const gpu_material *mats;
    const tex_desc *tex;
    const uint32_t *texels;
    bool disable_material;
    bool per_candidate = false;
    float xi = 0.0f;      // closest-hit rays: the trace's BLEND threshold
    uint32_t seed = 0u;   // shadow rays: per-candidate BLEND thresholds

    __device__ bool operator()(const float4 &rec, float bu, float bv) const
    {
        const gpu_material &m = mats[__float_as_uint(rec.z)];
        const uint32_t prim = __float_as_uint(rec.y);
        float a = 1.0f;
        if(!disable_material)
        {
            a = m.color.w;
            if(m.tex.x)
            {
                const float2 t0 = uv[idx[3 * prim]], t1 = uv[idx[3 * prim + 1]], t2 = uv[idx[3 * prim + 2]];
                const float b0 = 1.0f - bu - bv;
                const float u = (t0.x * b0 + t1.x * bu) + t2.x * bv, v = (t0.y * b0 + t1.y * bu) + t2.y * bv;
                a = a * sample_texture(tex[m.tex.x - 1u], texels, u, v, 0.0f).w;
            }
        }
        if(m.tex.y == VKRT_ALPHA_MASK) return !(a < __uint_as_float(m.tex.z));
        return !(a < (per_candidate ? rnd_peek(xxhash32(seed, prim)) : xi));
    }
};
__device__ __forceinline__ vec3 xyz(const float4 &f) { return vec3{f.x, f.y, f.z}; }
__device__ __forceinline__ float frac(float x) { return x - floorf(x); }

__global__ void __launch_bounds__(PT_BLOCK) k_pt_begin(pt_frame f, pt_buffers b, uint32_t sample)
{
    const uint32_t p = blockIdx.x * PT_BLOCK + threadIdx.x;
    // in the cache mode every train_stride-th pixel (from train_offset) traces a training path, queued on
    // its own; the others keep their order in the main queue
    const uint32_t n_train = f.cache_on ? f.n_train_paths : 0u;
    if(p == 0)
    {
        b.counter[0] = f.n_pix - n_train;
        b.counter[4] = 0u;
        b.counter[5] = 0u;
        b.counter[TRAIN_COUNTERS] = n_train;
    }
    if(p >= f.n_pix) return;
    const uint32_t x = p % f.width, y = p / f.width;
    uint4 st = b.state[p];
    float offset;
    if(sample == 0)
    {
        // rng_state = xxhash32(seed, W * y + x); sample_offset = 2 rnd - 1 (raypipeline.slang:95-100)
        uint32_t rng = xxhash32(f.seed, f.width * y + x);
        offset = 2.0f * rnd(rng) - 1.0f;
        st.x = rng;
        st.z = __float_as_uint(offset);
        b.acc[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    else
        offset = __uint_as_float(st.z);
    uint32_t rng = st.x;
    vec2 h = hammersley(sample, f.num_samples);
    float cx = static_cast<float>(x) + frac(h.x + offset);
    float cy = static_cast<float>(y) + frac(h.y + offset);
    vec2 xi;
    xi.x = rnd(rng);
    xi.y = rnd(rng);
    st.x = rng;
    st.y = 0u;// media index 0, no regularization, no caustic marker
    if(f.cache_on && p % f.train_stride == f.train_offset)
    {
        st.y = F_TRAIN;
        float4 *tv = b.tvtx + (size_t)(p / f.train_stride) * f.max_depth * VERTEX_F4;
        for(uint32_t d = 0; d < f.max_depth; ++d)
        {
            tv[d * VERTEX_F4 + 4] = make_float4(0.f, 0.f, 0.f, 0.f);
            tv[d * VERTEX_F4 + 5] = make_float4(0.f, 0.f, 0.f, 0.f);
            tv[d * VERTEX_F4 + 6] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    b.state[p] = st;

    // camera_ray (raypipeline.slang:58-89), perspective branch, in the camera's own basis
    float u = cx / static_cast<float>(f.width), v = cy / static_cast<float>(f.height);
    float dx = u * 2.0f - 1.0f, dy = v * 2.0f - 1.0f;
    vec2 disc = sample_unit_disc(xi);
    vec3 lens{f.aperture / 2.0f * disc.x, f.aperture / 2.0f * disc.y, 0.f};
    vec3 origin = f.eye + f.right * lens.x + f.up * lens.y;
    vec3 target = normalize3(vec3{dx * f.tan_half_x, -dy * f.tan_half_y, -1.0f});
    vec3 fpoint = target * (f.focal_distance / fabsf(target.z));
    vec3 dv = fpoint - lens;
    vec3 dir = normalize3(f.right * dv.x + f.up * dv.y - f.forward * dv.z);
    b.ray_o[p] = make_float4(origin.x, origin.y, origin.z, EPS);
    b.ray_d[p] = make_float4(dir.x, dir.y, dir.z, DEFAULT_TMAX);
    b.beta[p] = make_float4(1.f, 1.f, 1.f, 0.f);
    b.rad[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    b.iors[p] = make_float4(1.f, 1.f, 1.f, 1.f);
    if(f.any_volume) b.media[p] = make_uint2(0xffffffffu, 0xffffffffu);
    b.cone[p] = make_float2(f.cone_spread, 0.f);
    if(n_train == 0u) b.queue[0][p] = p;
    else if(st.y & F_TRAIN) b.train_queue[(p - f.train_offset) / f.train_stride] = p;
    else
    {
        const uint32_t below = p > f.train_offset ? (p - f.train_offset - 1u) / f.train_stride + 1u : 0u;
        b.queue[0][p - below] = p;
    }
}

__global__ void __launch_bounds__(PT_BLOCK) k_pt_zero(uint32_t *c) { c[threadIdx.x] = 0u; }

//! before a depth: the next queue, the shadow-ray count and the trace kernel's claim counters start at zero
//! (one thread per claim counter)
__global__ void k_pt_reset_depth(uint32_t *c, int next)
{
    if(threadIdx.x == 0)
    {
        c[next] = 0u;
        c[6] = 0u;
    }
    c[PT_CLAIMS + threadIdx.x * COUNTER_STRIDE] = 0u;
}

//! the closest hit of path p's current ray at this depth; back faces are culled while the path is in air.
//! BLEND candidates compare against rnd() of a copy of payload.rng_state = xxhash32(rng_state, depth)
//! (raypipeline.slang:150-152, :693, :716)
//! INST: an instanced scene, walked through its top level `tl` (the flat arguments are unused); the hit's
//! instance goes to b.hit_inst
template<int BLOCK = PT_BLOCK, class F = alpha_filter, bool INST = false>
__device__ __forceinline__ void trace_path(const vkrt_bvh_node *__restrict__ nodes, const float4 *__restrict__ tris,
                                           int32_t root, bool empty, const vkrt_bvh4_node *__restrict__ nodes4,
                                           int32_t root4, bool wide, const pt_buffers &b, uint32_t p, int32_t *smem,
                                           F flt, uint32_t depth, const pt_tlas &tl = pt_tlas{})
{
    float t = 0.f, u = 0.f, v = 0.f;
    uint32_t prim = VKRT_INVALID_PRIM;
    if constexpr(INST)
    {
        uint32_t inst = VKRT_INVALID_PRIM;
        if(!tl.empty)
        {
            const float4 ra = b.ray_o[p], rb = b.ray_d[p];
            const uint4 st = b.state[p];
            const bool in_air = (st.y & F_MEDIA_MASK) == 0u;
            if constexpr(F::active) flt.xi = rnd_peek(xxhash32(st.x, depth));
            t = rb.w;
            scene_walk<false, BLOCK, PT_LDS, F>(tl.tlas, tl.root, tl.order, tl.insts, tl.blases, ra, rb, smem, t, u, v,
                                                prim, inst, in_air, flt);
        }
        b.hit_inst[p] = inst;
    }
    else if(!empty)
    {
        ray_setup r = make_ray_setup(b.ray_o[p], b.ray_d[p]);
        const uint4 st = b.state[p];
        const bool in_air = (st.y & F_MEDIA_MASK) == 0u;
        if constexpr(F::active) flt.xi = rnd_peek(xxhash32(st.x, depth));
        // the shade kernel recovers HitKind (back face: negative determinant) from the vertices
        if(wide)
            traverse4<false, BLOCK, PT_LDS>(nodes4, tris, root4, r, smem, t, u, v, prim, in_air, flt);
        else
            traverse<false, BLOCK, PT_LDS>(nodes, tris, root, r, smem, t, u, v, prim, in_air, flt);
    }
    b.hit[p] = make_uint4(__float_as_uint(t), __float_as_uint(u), __float_as_uint(v), prim);
}

//! F is alpha_filter where some material is alpha-tested, keep_all elsewhere (the any-hit test and its texture
//! fetch compiled out)
template<class F, bool INST>
__global__ void __launch_bounds__(PT_TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(
        INST ? PT_TRACE_INST_WAVES_PER_SIMD : (F::active ? PT_TRACE_WAVES_PER_SIMD : PT_TRACE_PLAIN_WAVES_PER_SIMD))))
k_pt_trace(const vkrt_bvh_node *__restrict__ nodes, const float4 *__restrict__ tris, int32_t root, bool empty,
           const vkrt_bvh4_node *__restrict__ nodes4, int32_t root4, bool wide, pt_buffers b, int in, F flt,
           uint32_t depth, pt_tlas tl)
{
    __shared__ int32_t smem[PT_LDS * PT_TRACE_BLOCK];
    const uint32_t count = b.counter[in];
    const uint32_t tiles = (count + PT_TRACE_BLOCK - 1) / PT_TRACE_BLOCK;
    const uint32_t home = blockIdx.x % XCDS;
    uint32_t queue = 0;
    // one wave per workgroup: lane 0 claims the next tile of the queue and the wave reads it back; a grid of
    // one workgroup per tile or more (the full-grid launch) takes tile blockIdx.x and leaves
    const bool one_tile = gridDim.x >= tiles;
    bool done = false;
    while(!done)
    {
        uint32_t t = blockIdx.x;
        if(one_tile) done = true;
        else
        {
            if(threadIdx.x == 0) t = claim_tile(b.counter + PT_CLAIMS, PT_TRACE_CHUNK, home, tiles, queue);
            t = __builtin_amdgcn_readfirstlane(t);
        }
        if(t >= tiles) break;
        const uint32_t q = t * PT_TRACE_BLOCK + threadIdx.x;
        if(q < count)
            trace_path<PT_TRACE_BLOCK, F, INST>(nodes, tris, root, empty, nodes4, root4, wide, b, b.queue[in][q],
                                                smem + threadIdx.x, flt, depth, tl);
    }
}

//! eval_distant_lights (raypipeline.slang:728-744): sun discs seen by a ray that left the scene
__device__ vec3 eval_distant_lights(const vkrt_light *__restrict__ lights, uint32_t n, uint32_t mode, const vec3 &dir,
                                    float prev_bsdf_pdf)
{
    vec3 radiance{0.f, 0.f, 0.f};
    for(uint32_t i = 0; i < n; ++i)
    {
        const vkrt_light l = lights[i];
        if(l.type != VKRT_LIGHT_DIRECTIONAL) continue;
        const float pdf = light_pdf(l, vec3{0.f, 0.f, 0.f}, dir);
        if(pdf <= 0.0f) continue;
        const float w = mis_weight_hit(mode, prev_bsdf_pdf, pdf / static_cast<float>(n));
        radiance = radiance + (light_vec(l.color) * w) * l.intensity;
    }
    return radiance;
}

//! the ghost-light pass of raygen (raypipeline.slang:201-222): light bodies met inside the traced segment
//! (tmin, t_end) add their MIS-weighted radiance, scaled by the segment's throughput, to the path in light
//! order; `sum` receives the unscaled total for training records
__device__ vec3 ghost_lights(const vkrt_light *__restrict__ lights, const pt_frame &f, const vec3 &o, const vec3 &d,
                             float tmin, float t_end, const vec3 &seg_beta, float prev_pdf, bool caustic,
                             vec3 radiance, vec3 &sum)
{
    sum = vec3{0.f, 0.f, 0.f};
    for(uint32_t li = 0; li < f.n_lights; ++li)
    {
        const vkrt_light l = lights[li];
        const light_hit hit = intersect_light(l, o, d);
        if(hit.t <= tmin || hit.t >= t_end) continue;
        if(f.suppress_caustics && caustic) continue;
        const float p_nee = light_pdf(l, o, d) / static_cast<float>(f.n_lights);
        const float w = mis_weight_hit(f.mis_mode, prev_pdf, p_nee);
        radiance = radiance + mul3(seg_beta * w, light_vec(l.color)) * l.intensity;
        sum = sum + (light_vec(l.color) * w) * l.intensity;
    }
    return radiance;
}

constexpr uint32_t NO_MEDIUM = 0xffffu;

__device__ __forceinline__ uint32_t media_level(const uint2 &m, uint32_t level)
{
    return ((level < 2u ? m.x : m.y) >> (16u * (level & 1u))) & 0xffffu;
}

__device__ __forceinline__ void set_media_level(uint2 &m, uint32_t level, uint32_t mid)
{
    uint32_t &w = level < 2u ? m.x : m.y;
    const uint32_t sh = 16u * (level & 1u);
    w = (w & ~(0xffffu << sh)) | ((mid & 0xffffu) << sh);
}

//! the participating medium of the segment that ended at a hit (raypipeline.slang:380-415), absorption
//! only (a medium of single-scattering albedo 0): a channel is picked, a collision distance drawn against its
//! extinction; a collision before the surface absorbs the path (the reference's throughput becomes 0 there),
//! else the throughput takes the segment's transmittance over the pdf of passing. Returns true on a collision
__device__ __forceinline__ bool medium_collision(const float4 &sigma, float t_hit, vec3 &beta, uint32_t &rng)
{
    if(!(sigma.x > 0.f || sigma.y > 0.f || sigma.z > 0.f)) return false;
    const int channel = static_cast<int>(fminf(rnd(rng) * 3.0f, 2.0f));
    const float sc = channel == 0 ? sigma.x : (channel == 1 ? sigma.y : sigma.z);
    const float t = fminf(-logf(1.0f - rnd(rng)) / fmaxf(sc, EPS), t_hit);
    if(t < t_hit) return true;
    const vec3 tr{expf(-sigma.x * t), expf(-sigma.y * t), expf(-sigma.z * t)};
    const float pdf = fmaxf((tr.x + tr.y + tr.z) / 3.0f, PDF_EPS);
    beta = vec3{beta.x * (tr.x / pdf), beta.y * (tr.y / pdf), beta.z * (tr.z / pdf)};
    return false;
}

//! one path's vertex at `depth` (closest hit or miss): emission, next-event sample, BSDF sample, Russian
//! roulette, light bodies met on the segment; writes the path state back and reports what the path needs
//! next. LIGHTS (the frame has analytic lights) compiles next-event estimation and the light-body passes in,
//! so the unlit kernels keep their smaller register footprint
template<bool LIGHTS, bool INST = false>
__device__ __forceinline__ void shade_path(const pt_frame &f, const pt_buffers &b, const pt_scene_dev &s,
                                           uint32_t depth, uint32_t p, bool &alive, bool &to_cache, bool &need_shadow,
                                           float4 (&shadow_rec)[SHADOW_F4], float4 (&qfeat)[QUERY_F4])
{
    uint4 st = b.state[p];
    // rng_state = xxhash32(rng_state, i) in the raygen chain (raypipeline.slang:150); the hit program
    // draws from a copy
    st.x = xxhash32(st.x, depth);
    uint32_t rng = st.x;
    uint32_t flags = st.y;
    float4 ro = b.ray_o[p], rd = b.ray_d[p];
    vec3 dir = xyz(rd);
    const float4 beta4 = b.beta[p];
    vec3 beta = xyz(beta4);
    float prev_pdf = beta4.w;// payload.prev_bsdf_pdf
    // the traced segment, for the ghost-light pass after shading
    const vec3 seg_o = xyz(ro), seg_d = dir, seg_beta = beta;
    const float seg_tmin = ro.w, seg_tmax = rd.w, seg_prev_pdf = prev_pdf;
    const bool seg_caustic = (flags & F_CAUSTIC) != 0u;
    vec3 radiance = xyz(b.rad[p]);
    float2 cone = b.cone[p];// ray cone: spread angle, width (half precision values)
    float4 iors = b.iors[p];
    float ior_stack[MEDIA_STACK] = {iors.x, iors.y, iors.z, iors.w};
    uint32_t media_index = flags & F_MEDIA_MASK;
    uint2 media = f.any_volume ? b.media[p] : make_uint2(0xffffffffu, 0xffffffffu);
    uint4 h = b.hit[p];
    bool stop = false;
    to_cache = false;
    uint32_t object = NO_OBJECT;
    vec3 position{0.f, 0.f, 0.f}, normal{0.f, 0.f, 0.f}, albedo{0.f, 0.f, 0.f};
    uint32_t media_op = 0;// 0 NoOp, 1 Enter, 2 Leave
    float next_ior = 1.f;
    const bool train = (flags & F_TRAIN) != 0u;
    // training vertex: kind 1 = surface (features, emission, throughput weight), 2 = miss (sky radiance)
    uint32_t vkind = 0;
    vec3 v_e{0.f, 0.f, 0.f}, v_w{0.f, 0.f, 0.f};
    float4 feat[4];

    if(h.w == VKRT_INVALID_PRIM)
    {
        // miss_default (raypipeline.slang:746-761)
        stop = true;
        vec3 col = environment_white(dir);
        radiance = radiance + mul3(beta * f.environment, col);
        vkind = 2;
        v_e = col * f.environment;
        if(LIGHTS && !(f.suppress_caustics && seg_caustic))
        {
            const vec3 dl = eval_distant_lights(s.lights, f.n_lights, f.mis_mode, dir, prev_pdf);
            radiance = radiance + mul3(beta, dl);
            v_e = v_e + dl;
        }
    }
    else if(f.any_volume &&
            medium_collision(media_level(media, media_index) == NO_MEDIUM
                                     ? make_float4(0.f, 0.f, 0.f, 0.f)
                                     : s.mats[media_level(media, media_index)].vol,
                             __uint_as_float(h.x), beta, rng))
    {
        // absorbed in the medium before the surface: the path ends with what it has gathered
        stop = true;
    }
    else
    {
        const uint32_t prim = h.w;
        const float u = __uint_as_float(h.y), v = __uint_as_float(h.z);
        const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
        const vec3 v0 = ld3f(s.pos, i0), v1 = ld3f(s.pos, i1), v2 = ld3f(s.pos, i2);
        const vec3 e1 = v1 - v0, e2 = v2 - v0;
        // HitKind: the determinant's sign in the triangle's own space (an instance's ray direction turned by
        // its world-to-object rows, as the walk turned it)
        vec3 dir_o = dir;
        const float4 *xf = nullptr;
        if constexpr(INST)
        {
            const uint32_t inst = b.hit_inst[p];
            xf = s.inst_xf + 3 * (size_t)inst;
            const inst_dev I = s.insts[inst];
            dir_o = vec3{(I.w0.x * dir.x + I.w0.y * dir.y) + I.w0.z * dir.z,
                         (I.w1.x * dir.x + I.w1.y * dir.y) + I.w1.z * dir.z,
                         (I.w2.x * dir.x + I.w2.y * dir.y) + I.w2.z * dir.z};
        }
        const bool backface = dot(e1, vierkant::rt::cross(dir_o, e2)) < 0.f;
        // interpolate_vertex (raypipeline.slang:336-358): object-space interpolation, then the instance's
        // transform on the position and its rotation on the normal (identity for a flat scene)
        const float bx = 1.0f - u - v, by = u, bz = v;
        position = (v0 * bx + v1 * by) + v2 * bz;
        vec3 n = (ld3f(s.nrm, i0) * bx + ld3f(s.nrm, i1) * by) + ld3f(s.nrm, i2) * bz;
        normal = normalize3(n);
        if constexpr(INST)
        {
            position = apply_instance(xf, position);
            normal = rotate_quat(xf[0], normal);
        }
        if(backface) normal = -normal;
        object = s.tri_object[prim];
        const gpu_material gm = s.mats[s.object_material[object]];
        // tmax_inside_volume = 2 * length(scale * aabb.half_extents()) (raypipeline.slang:370)
        float tmax_inside;
        if constexpr(INST)
        {
            const vec3 he = ld3f(s.object_he, object);
            const float4 ts = xf[1], s2 = xf[2];
            const float hx = ts.w * he.x, hy = s2.x * he.y, hz = s2.y * he.z;
            tmax_inside = 2.f * sqrtf((hx * hx + hy * hy) + hz * hz);
        }
        else
            tmax_inside = s.object_tmax[object];
        const vec3 V = -dir;
        ro = make_float4(position.x, position.y, position.z, ro.w);

        material m;
        m.color = xyz(gm.color);
        m.alpha = gm.color.w;
        m.emission = xyz(gm.emission);
        m.emission_strength = gm.emission.w;
        m.metalness = gm.mrti.x;
        m.roughness = gm.mrti.y;
        m.transmission = gm.mrti.z;
        m.ior = gm.mrti.w;
        m.clearcoat = gm.cc.x;
        m.clearcoat_roughness = gm.cc.y;
        m.sheen_roughness = gm.cc.z;
        m.diffuse_transmission = gm.cc.w;
        m.sheen_color = xyz(gm.sheen);
        m.diffuse_transmission_color = xyz(gm.dtc);
        if(gm.tex.x | gm.tex2.x | gm.tex2.y | gm.tex2.z | gm.tex2.w | gm.tex3.x | gm.tex3.y)
        {
            // every texture at the ray cone's level of detail (raypipeline.slang:300-321, 368-375, 526-606):
            // lambda = 0.5 log2(uv area / world area) + log2|cone width| + 0.5 log2(w h) - log2(N.V), with the
            // width from before this segment's propagation
            const float2 t0 = s.uv[i0], t1 = s.uv[i1], t2 = s.uv[i2];
            const float tu = (t0.x * bx + t1.x * by) + t2.x * bz, tv = (t0.y * bx + t1.y * by) + t2.y * bz;
            vec3 w0 = v0, w1 = v1, w2 = v2;
            if constexpr(INST)
            {
                // lod_constant: the triangle's area in world space
                w0 = apply_instance(xf, v0);
                w1 = apply_instance(xf, v1);
                w2 = apply_instance(xf, v2);
            }
            const float p_a = length3(vierkant::rt::cross(w1 - w0, w2 - w0));
            const float t_a = fabsf((t1.x - t0.x) * (t2.y - t0.y) - (t2.x - t0.x) * (t1.y - t0.y));
            const float lod_cone = 0.5f * log2f(t_a / p_a) + log2f(fabsf(cone.y));
            const float lod_nov = log2f(fabsf(dot(V, normal)));
            auto lookup = [&](uint32_t slot) {
                const tex_desc &td = s.tex[slot - 1u];
                float lambda = lod_cone;
                lambda += 0.5f * log2f(static_cast<float>(td.width) * static_cast<float>(td.height));
                lambda -= lod_nov;
                return sample_texture(td, s.texels, tu, tv, lambda);
            };
            if(gm.tex.x)
            {
                const float4 c = lookup(gm.tex.x);
                m.color = vec3{m.color.x * c.x, m.color.y * c.y, m.color.z * c.z};
                m.alpha = m.alpha * c.w;
            }
            if(gm.tex2.x)
            {
                // tangent-space normal map, bitangent handedness flipped on back faces, hemisphere clamp
                const float4 c = lookup(gm.tex2.x);
                vec3 nt;
                nt.x = 2.0f * (c.x - 0.5f);
                nt.y = 2.0f * (c.y - 0.5f);
                nt.z = sqrtf(fmaxf(0.0f, 1.0f - nt.x * nt.x - nt.y * nt.y));
                vec3 tg = normalize3((ld3f(s.tan, i0) * bx + ld3f(s.tan, i1) * by) + ld3f(s.tan, i2) * bz);
                if constexpr(INST) tg = rotate_quat(xf[0], tg);
                vec3 bt = normalize3(vierkant::rt::cross(normal, tg));
                if(backface) bt = -bt;
                const vec3 nm = normalize3((tg * nt.x + bt * nt.y) + normal * nt.z);
                normal = normalize3(nm + normal * fmaxf(0.0f, -dot(nm, normal) + 1e-3f));
            }
            if(gm.tex2.y)
            {
                const float4 c = lookup(gm.tex2.y);
                if(m.emission.x > 0.0f || m.emission.y > 0.0f || m.emission.z > 0.0f)
                    m.emission = vec3{m.emission.x * c.x, m.emission.y * c.y, m.emission.z * c.z};
                else
                    m.emission = vec3{c.x, c.y, c.z};
            }
            if(gm.tex2.z)
            {
                const float4 c = lookup(gm.tex2.z);
                m.roughness = m.roughness * c.y;
                m.metalness = m.metalness * c.z;
            }
            if(gm.tex2.w) m.transmission = m.transmission * lookup(gm.tex2.w).x;
            if(gm.tex3.x) m.diffuse_transmission = m.diffuse_transmission * lookup(gm.tex3.x).w;
            if(gm.tex3.y)
            {
                const float4 c = lookup(gm.tex3.y);
                m.diffuse_transmission_color = vec3{m.diffuse_transmission_color.x * c.x,
                                                    m.diffuse_transmission_color.y * c.y,
                                                    m.diffuse_transmission_color.z * c.z};
            }
        }
        if(f.disable_material)
        {
            m.color = splat(0.8f);
            m.alpha = 1.0f;
        }
        // propagate(cone, 0, RayTCurrent()) (raypipeline.slang:280-286, :608), stored as halves
        cone.y = half_value(cone.y + cone.x * __uint_as_float(h.x));

        feat[0] = make_float4(position.x, position.y, position.z, m.roughness);
        feat[1] = make_float4(V.x, V.y, V.z, m.metalness);
        feat[2] = make_float4(normal.x, normal.y, normal.z, m.transmission);
        feat[3] = make_float4(m.color.x, m.color.y, m.color.z, 0.f);
        if(f.cache_on && !train && depth == f.cache_depth)
        {
            // terminate into the cache: the query kernel adds beta * L_o(x, wo), emission included
            stop = true;
            to_cache = true;
        }
        else
        {
        const float media_ior = ior_stack[media_index];
        const float last_ior = ior_stack[media_index > 0u ? media_index - 1u : 0u];
        float eta = backface ? m.ior / last_ior : media_ior / m.ior;
        eta += EPS;
        next_ior = m.ior;
        // decided on the material's own transmission (before its texture) or its interior medium
        // (raypipeline.slang:517-522)
        const bool media_boundary = gm.mrti.z > 0.0f || gm.vol.w > 0.0f;
        media_op = !media_boundary ? 0u : (backface ? 2u : 1u);
        if(depth == 0) albedo = m.color;
        const vec3 sn = normal;
        const vec3 ff = dot(dir, sn) < 0.f ? sn : -sn;// faceforward(sn, ray dir, sn)
        vec3 emission = m.emission * (dot(sn, ff) > 0.f ? m.emission_strength : 0.0f);
        radiance = radiance + mul3(beta, emission);
        vkind = 1;
        v_e = emission;
        if(flags & F_REGULARIZE) m.roughness = fmaxf(m.roughness, REGULARIZE_MIN_ROUGHNESS);
        bsdf_sample bs = sample_disney(m, sn, V, eta, rng);
        if(LIGHTS)
        {
            // nee_surface (direct_lighting.slang:296-336): one light picked uniformly, one sample of it; the
            // shadow ray's transmittance is applied by k_pt_shadow
            const uint32_t nl = f.n_lights;
            const uint32_t index = nl > 1u ? min(static_cast<uint32_t>(rnd(rng) * static_cast<float>(nl)), nl - 1u)
                                            : 0u;
            vec2 xi;
            xi.x = rnd(rng);
            xi.y = rnd(rng);
            const vkrt_light l = s.lights[index];
            const light_sample ls = sample_light(l, position, xi);
            if(ls.radiance.x > 0.0f || ls.radiance.y > 0.0f || ls.radiance.z > 0.0f)
            {
                float bsdf_pdf = 0.0f;
                const float cos_theta = fmaxf(dot(sn, ls.direction), 0.0f);
                const vec3 F = eval_disney(m, ls.direction, sn, V, eta, bsdf_pdf);
                const float weight = bsdf_reachable(l)
                                             ? mis_weight_nee(f.mis_mode, ls.pdf / static_cast<float>(nl), bsdf_pdf)
                                             : 1.0f;
                const vec3 org = position + sn * EPS;
                shadow_rec[0] = make_float4(org.x, org.y, org.z, 0.0f);
                shadow_rec[1] = make_float4(ls.direction.x, ls.direction.y, ls.direction.z, ls.dist);
                shadow_rec[2] = make_float4(ls.radiance.x, ls.radiance.y, ls.radiance.z, weight);
                shadow_rec[3] = make_float4(F.x, F.y, F.z, cos_theta);
                shadow_rec[4] = make_float4(beta.x, beta.y, beta.z, ls.pdf);
                const uint32_t vtx = train ? (p / f.train_stride) * f.max_depth + depth : NO_VERTEX;
                // the state the transmittance test would continue drawing from: BLEND candidates' seed
                shadow_rec[5] = make_float4(__uint_as_float(p), __uint_as_float(vtx), __uint_as_float(rng), 0.f);
                need_shadow = true;
            }
        }
        if(!(bs.pdf > 0.0f)) { stop = true; }
        else
        {
            prev_pdf = LIGHTS ? bs.pdf : 0.0f;
            dir = bs.direction;
            const bool specular = bs.pdf < REGULARIZE_PDF_THRESHOLD;
            if(specular) flags |= F_REGULARIZE;
            const bool caustic = specular && ((flags & F_CAUSTIC) || bs.transmission);
            flags = caustic ? (flags | F_CAUSTIC) : (flags & ~F_CAUSTIC);
            const float cos_theta = fabsf(dot(sn, bs.direction));
            const float denom = fmaxf(bs.pdf, PDF_EPS);
            vec3 thr = bs.F * cos_theta;
            thr = vec3{thr.x / denom, thr.y / denom, thr.z / denom};
            if(depth >= 1 && f.max_path_beta > 0.0f)
                thr = vec3{fminf(thr.x, f.max_path_beta), fminf(thr.y, f.max_path_beta),
                           fminf(thr.z, f.max_path_beta)};
            beta = mul3(beta, thr);
            v_w = thr;
            if(bs.transmission) rd.w = media_op == 1u ? tmax_inside : DEFAULT_TMAX;
            media_op = bs.transmission ? media_op : 0u;
            // Russian roulette (raypipeline.slang:667-677)
            if(max3(beta) <= 0.05f * static_cast<float>(depth))
            {
                float qq = 1.0f - max3(beta);
                if(rnd(rng) < qq)
                {
                    stop = true;
                    v_w = vec3{0.f, 0.f, 0.f};
                }
                else
                {
                    const float keep = 1.0f - qq;
                    beta = vec3{beta.x / keep, beta.y / keep, beta.z / keep};
                    v_w = vec3{v_w.x / keep, v_w.y / keep, v_w.z / keep};
                }
            }
        }
        }
    }
    vec3 ghost{0.f, 0.f, 0.f};
    if(LIGHTS)
    {
        // segment end: the hit point, or the traced tmax on a miss
        const float t_end = h.w == VKRT_INVALID_PRIM ? seg_tmax : length3(position - seg_o);
        radiance = ghost_lights(s.lights, f, seg_o, seg_d, seg_tmin, t_end, seg_beta, seg_prev_pdf, seg_caustic,
                                radiance, ghost);
    }
    if(train && vkind != 0u)
    {
        float4 *tv = b.tvtx + ((size_t)(p / f.train_stride) * f.max_depth + depth) * VERTEX_F4;
        tv[6] = make_float4(ghost.x, ghost.y, ghost.z, 0.f);
        if(vkind == 1u)
        {
            tv[0] = feat[0];
            tv[1] = feat[1];
            tv[2] = feat[2];
            tv[3] = feat[3];
        }
        tv[4] = make_float4(v_e.x, v_e.y, v_e.z, __uint_as_float(vkind));
        tv[5] = make_float4(v_w.x, v_w.y, v_w.z, 0.f);
    }
    if(to_cache)
    {
        qfeat[0] = feat[0];
        qfeat[1] = feat[1];
        qfeat[2] = feat[2];
        qfeat[3] = feat[3];
        qfeat[4] = make_float4(__uint_as_float(p), beta.x, beta.y, beta.z);
    }

    // media stack update in the raygen loop (raypipeline.slang:195-200)
    if(media_op == 1u)
    {
        media_index = min(media_index + 1u, MEDIA_STACK - 1u);
        ior_stack[media_index] = round_half(next_ior);
        if(f.any_volume) set_media_level(media, media_index, s.object_material[object]);
    }
    else if(media_op == 2u)
        media_index = media_index > 0u ? media_index - 1u : 0u;
    flags = (flags & ~F_MEDIA_MASK) | media_index;

    if(depth == 0)
    {
        const bool miss = object == NO_OBJECT;
        if(miss) flags |= F_FIRST_MISS;
        else flags &= ~F_FIRST_MISS;
        const float zdepth = miss ? 0.0f : dot(position - f.eye, f.forward);
        b.info[p] = make_uint4(__float_as_uint(zdepth), pack_normal_oct(normal), object, pack_rgba8(albedo, 1.0f));
    }
    st.y = flags;
    b.state[p] = st;
    alive = !stop && depth + 1 < f.max_depth;
    if(alive)
    {
        b.ray_o[p] = ro;
        b.ray_d[p] = make_float4(dir.x, dir.y, dir.z, rd.w);
        b.beta[p] = make_float4(beta.x, beta.y, beta.z, prev_pdf);
        b.rad[p] = make_float4(radiance.x, radiance.y, radiance.z, 0.f);
        b.cone[p] = cone;
        b.iors[p] = make_float4(ior_stack[0], ior_stack[1], ior_stack[2], ior_stack[3]);
        if(f.any_volume) b.media[p] = media;
    }
    else
    {
        // result.radiance += payload.radiance; coverage = max(coverage, first hit ? 1 : 0)
        float4 a = b.acc[p];
        const float cov = (flags & F_FIRST_MISS) ? 0.0f : 1.0f;
        b.acc[p] = make_float4(a.x + radiance.x, a.y + radiance.y, a.z + radiance.z, fmaxf(a.w, cov));
    }
}

//! queue entry q of a depth (inactive past the count): shade_path, then survivors, cache queries and shadow
//! rays are appended with one ballot and one atomic per wave each
template<bool LIGHTS, bool INST>
__device__ __forceinline__ void shade_queue_entry(const pt_frame &f, const pt_buffers &b, const pt_scene_dev &s,
                                                  uint32_t depth, int in, uint32_t q, uint32_t count)
{
    const bool active = q < count;
    bool alive = false;
    bool to_cache = false;
    bool need_shadow = false;
    float4 shadow_rec[SHADOW_F4];
    float4 qfeat[QUERY_F4];
    uint32_t p = 0;
    if(active)
    {
        p = b.queue[in][q];
        shade_path<LIGHTS, INST>(f, b, s, depth, p, alive, to_cache, need_shadow, shadow_rec, qfeat);
    }
    // append survivors to the next queue: one ballot and one atomic per wave
    const uint64_t mask = __ballot(alive);
    const uint32_t lane = __lane_id();
    const uint64_t n_active = __ballot(active);
    uint32_t base = 0;
    if(lane == 0)
    {
        if(mask) base = atomicAdd(&b.counter[in ^ 1], static_cast<uint32_t>(__popcll(mask)));
        if(n_active) atomicAdd(reinterpret_cast<unsigned long long *>(&b.counter[2]), (unsigned long long)__popcll(n_active));
    }
    base = __shfl(base, 0);
    if(alive)
    {
        const uint64_t below = lane == 0 ? 0ull : (mask & (~0ull >> (64 - lane)));
        b.queue[in ^ 1][base + __popcll(below)] = p;
    }
    // cache queries: one ballot and one atomic per wave
    const uint64_t qmask = __ballot(to_cache);
    if(qmask)
    {
        uint32_t qbase = 0;
        if(lane == 0) qbase = atomicAdd(&b.counter[4], static_cast<uint32_t>(__popcll(qmask)));
        qbase = __shfl(qbase, 0);
        if(to_cache)
        {
            const uint64_t below = lane == 0 ? 0ull : (qmask & (~0ull >> (64 - lane)));
            float4 *dst = b.qrec + (size_t)(qbase + __popcll(below)) * QUERY_F4;
            for(uint32_t k = 0; k < QUERY_F4; ++k) dst[k] = qfeat[k];
        }
    }
    // next-event shadow rays: one ballot and one atomic per wave
    const uint64_t smask = __ballot(need_shadow);
    if(smask)
    {
        uint32_t sbase = 0;
        if(lane == 0) sbase = atomicAdd(&b.counter[6], static_cast<uint32_t>(__popcll(smask)));
        sbase = __shfl(sbase, 0);
        if(need_shadow)
        {
            const uint64_t below = lane == 0 ? 0ull : (smask & (~0ull >> (64 - lane)));
            float4 *dst = b.shadow + (size_t)(sbase + __popcll(below)) * SHADOW_F4;
            for(uint32_t k = 0; k < SHADOW_F4; ++k) dst[k] = shadow_rec[k];
        }
    }
}

//! one depth of shading over a compacted queue; survivors, cache queries and shadow rays are appended with
//! one ballot and one atomic per wave each
//! (one tile per workgroup over a grid that covers every pixel: a grid-striding loop here keeps more values
//! live across iterations, 190 VGPRs against 161, and drops the kernel from 3 waves per SIMD to 2)
template<bool LIGHTS, bool INST>
__global__ void __launch_bounds__(PT_BLOCK) k_pt_shade(pt_frame f, pt_buffers b, pt_scene_dev s, uint32_t depth,
                                                       int in)
{
    const uint32_t count = b.counter[in];
    const uint32_t first = blockIdx.x * PT_BLOCK;
    if(first < count) shade_queue_entry<LIGHTS, INST>(f, b, s, depth, in, first + threadIdx.x, count);
}

//! training targets: L_d = e_d + w_d * (g_{d+1} + L_{d+1}) from the last vertex back, where e_d holds the
//! vertex's emission and next-event estimate (a miss slot: the sky and sun discs) and g_{d+1} the light bodies
//! met on the segment leaving it; every surface vertex becomes a training record, at most train_max per sample
__global__ void __launch_bounds__(PT_BLOCK) k_pt_train_targets(pt_frame f, pt_buffers b)
{
    const uint32_t t = blockIdx.x * PT_BLOCK + threadIdx.x;
    if(t >= f.n_train_paths) return;
    if((size_t)t * f.train_stride + f.train_offset >= f.n_pix) return;
    const float4 *tv = b.tvtx + (size_t)t * f.max_depth * VERTEX_F4;
    vec3 Lin{0.f, 0.f, 0.f};
    for(int d = (int)f.max_depth - 1; d >= 0; --d)
    {
        const float4 e = tv[d * VERTEX_F4 + 4], w = tv[d * VERTEX_F4 + 5], g = tv[d * VERTEX_F4 + 6];
        const uint32_t kind = __float_as_uint(e.w);
        if(kind == 0u) continue;
        const vec3 L{e.x + w.x * Lin.x, e.y + w.y * Lin.y, e.z + w.z * Lin.z};
        Lin = vec3{g.x + L.x, g.y + L.y, g.z + L.z};
        if(kind != 1u) continue;
        const uint32_t k = atomicAdd(&b.counter[5], 1u);
        if(k >= f.train_max) continue;
        float4 *r = b.trec + (size_t)k * 4;
        r[0] = tv[d * VERTEX_F4 + 0];
        r[1] = tv[d * VERTEX_F4 + 1];
        r[2] = tv[d * VERTEX_F4 + 2];
        r[3] = tv[d * VERTEX_F4 + 3];
        b.ttgt[k] = make_float4(L.x, L.y, L.z, 0.f);
    }
}

__global__ void __launch_bounds__(PT_BLOCK) k_pt_resolve(pt_frame f, pt_buffers b)
{
    const uint32_t p = blockIdx.x * PT_BLOCK + threadIdx.x;
    if(p >= f.n_pix) return;
    float4 a = b.acc[p];
    const float ns = static_cast<float>(f.num_samples);
    vec3 r{a.x / ns, a.y / ns, a.z / ns};
    const float coverage = a.w;
    vkrt_pixel out;
    vec3 accum = r;
    if(f.batch_index != 0)
    {
        // running mean over batches (raypipeline.slang:258-266)
        const float bs = static_cast<float>(f.batch_index);
        const vkrt_pixel old = b.pixels[p];
        vec3 sum{accum.x + bs * old.radiance[0], accum.y + bs * old.radiance[1], accum.z + bs * old.radiance[2]};
        const float div = static_cast<float>(f.batch_index + 1);
        accum = vec3{sum.x / div, sum.y / div, sum.z / div};
    }
    const uint4 info = b.info[p];
    out.radiance[0] = accum.x;
    out.radiance[1] = accum.y;
    out.radiance[2] = accum.z;
    out.alpha = coverage;// the stored alpha is this batch's coverage (raypipeline.slang:268-269)
    out.depth = __uint_as_float(info.x);
    out.normal = info.y;
    out.object_id = info.z;
    out.albedo = info.w;
    b.pixels[p] = out;
}

// ---------------------------------------------------------------------------------------------
// transmittance shadow rays (transmittance_test, shaders/slang/ray/visibility_test.slang:94-165): any hit of
// an opaque triangle (transmission == 0; RayBuilder.cpp:212-214 marks only those geometries opaque) in
// (tmin, tmax) gives 0; otherwise every transmissive hit multiplies the transmittance by transmission * color.
// The reference multiplies in the driver's candidate order, whose rounding is unspecified; here the factors are
// multiplied in ascending (t, triangle) order, so the result is defined bit for bit: each pass walks all leaves
// whose boxes meet [tmin, tmax], keeps the TK nearest candidates above the previous pass's last one in a
// sorted register list, and another pass runs only when more remained.
// ---------------------------------------------------------------------------------------------
constexpr int TK = 8;
constexpr int TR_STACK = 64;

//! the transmittance of one ray through the scene: zero when an opaque triangle lies in (tmin, tmax),
//! else the product of transmission * color over the transmissive hits in ascending (t, triangle) order
template<class F = alpha_filter>
__device__ float4 transmittance_walk(const vkrt_bvh_node *__restrict__ nodes2, int32_t root2,
                                     const vkrt_bvh4_node *__restrict__ nodes4, int32_t root4, bool wide, bool empty,
                                     const float4 *__restrict__ tris, const gpu_material *__restrict__ mats,
                                     const ray_setup &r, bool cull, int32_t *lds, const F &flt)
{
    int32_t spill[TR_STACK - PT_LDS];
    lane_stack<PT_BLOCK, PT_LDS> st{lds, spill};
    float T0 = 1.f, T1 = 1.f, T2 = 1.f, D0 = 0.f, D1 = 0.f, D2 = 0.f;
    bool blocked = false, more = !empty, first_pass = true;
    float lb_t = 0.f;
    uint32_t lb_p = 0;
    while(more && !blocked)
    {
        float kt[TK];
        uint32_t kp[TK], km[TK];
#pragma unroll
        for(int j = 0; j < TK; ++j)
        {
            kt[j] = __int_as_float(0x7f800000);
            kp[j] = 0xffffffffu;
            km[j] = 0u;
        }
        int cnt = 0;
        more = false;
        // the running bound of the pass: once TK candidates are kept, boxes beyond the farthest of them can add
        // nothing to this pass (whatever lies there, an opaque blocker included, is the next pass's); a later
        // pass skips boxes that end before the previous pass's last candidate (an opaque blocker there would
        // have ended the walk already)
        float bound = r.tmax;
        const float lo = first_pass ? r.tmin : fmaxf(r.tmin, lb_t);
        int sp = 0;
        int32_t ref = wide ? root4 : root2;
        while(true)
        {
            if(ref >= 0)
            {
                int32_t c[4];
                int nc;
                if(wide) nc = visit4(nodes4, ref, r, bound, c, lo);
                else
                {
                    const vkrt_bvh_node nd = nodes2[ref];
                    float tn0, tf0, tn1, tf1;
                    slab_interval(r, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.c.x, nd.c.y, tn0, tf0);
                    slab_interval(r, nd.b.x, nd.b.y, nd.b.z, nd.b.w, nd.c.z, nd.c.w, tn1, tf1);
                    const bool h0 = tn0 <= tf0 && tf0 >= lo && tn0 <= bound;
                    const bool h1 = tn1 <= tf1 && tf1 >= lo && tn1 <= bound;
                    nc = 0;
                    if(h0) c[nc++] = nd.d.x;
                    if(h1) c[nc++] = nd.d.y;
                }
                if(nc)
                {
                    for(int j = nc - 1; j >= 1; --j) st.put(sp++, c[j]);
                    ref = c[0];
                    continue;
                }
            }
            else
            {
                uint32_t first, count;
                leaf_decode(ref, first, count);
                for(uint32_t k = 0; k < count; ++k)
                {
                    const float4 *tr = tris + 3 * (size_t)(first + k);
                    const float4 p0 = tr[0], p1 = tr[1], p2 = tr[2];
                    const bool two_sided = (__float_as_uint(p2.w) & TRI_TWO_SIDED) != 0u;
                    float t, u, v;
                    if(!vierkant::rt::mt_fast(r.o, r.d, vec3{p0.x, p0.y, p0.z}, vec3{p0.w, p1.x, p1.y},
                                              vec3{p1.z, p1.w, p2.x}, t, u, v, cull && !two_sided) ||
                       !(t > r.tmin && t < r.tmax))
                        continue;
                    const uint32_t mid = __float_as_uint(p2.z);
                    const gpu_material m = mats[mid];
                    if(!(m.mrti.z > 0.0f))
                    {
                        // alpha-cutout surfaces let the shadow ray through (visibility_test.slang:145-150)
                        if(filtered_out(flt, p2, u, v)) continue;
                        blocked = true;
                        break;
                    }
                    const uint32_t prim = __float_as_uint(p2.y);
                    const uint32_t tag = candidate_tag(mid, m, r.d, p0, p1, p2);
                    if(!first_pass && (t < lb_t || (t == lb_t && prim <= lb_p))) continue;
                    if(cnt == TK && (t > kt[TK - 1] || (t == kt[TK - 1] && prim > kp[TK - 1])))
                    {
                        more = true;
                        continue;
                    }
                    if(cnt == TK) more = true;// the largest kept candidate is pushed out
                    // insert into the sorted list (ascending t, then triangle)
                    float ct = t;
                    uint32_t cp = prim, cm = tag;
#pragma unroll
                    for(int j = 0; j < TK; ++j)
                    {
                        const bool less = ct < kt[j] || (ct == kt[j] && cp < kp[j]);
                        const float tt = kt[j];
                        const uint32_t tp = kp[j], tm = km[j];
                        kt[j] = less ? ct : tt;
                        kp[j] = less ? cp : tp;
                        km[j] = less ? cm : tm;
                        ct = less ? tt : ct;
                        cp = less ? tp : cp;
                        cm = less ? tm : cm;
                    }
                    cnt = cnt < TK ? cnt + 1 : TK;
                    if(cnt == TK)
                    {
                        bound = kt[TK - 1];
                        more = true;
                    }
                }
                if(blocked) break;
            }
            if(sp == 0) break;
            --sp;
            ref = st.get(sp);
        }
        if(blocked) break;
        // multiply in ascending (t, triangle) order; the triangle's material comes from its leaf record
#pragma unroll
        for(int j = 0; j < TK; ++j)
        {
            if(j < cnt)
            {
                const gpu_material m = mats[km[j] & ~KM_FRONT];
                T0 = T0 * (m.mrti.z * m.color.x);
                T1 = T1 * (m.mrti.z * m.color.y);
                T2 = T2 * (m.mrti.z * m.color.z);
                if(m.vol.w > 0.0f)
                {
                    const bool front = (km[j] & KM_FRONT) != 0u;
                    D0 += front ? -m.vol.x * kt[j] : m.vol.x * kt[j];
                    D1 += front ? -m.vol.y * kt[j] : m.vol.y * kt[j];
                    D2 += front ? -m.vol.z * kt[j] : m.vol.z * kt[j];
                }
                lb_t = kt[j];
                lb_p = kp[j];
            }
        }
        first_pass = false;
    }
    // unpaired entries (the light inside a volume) must not amplify: the depth is clamped at 0
    // (exp(-0) = 1 exactly: scenes without volumes skip the exp)
    return blocked ? make_float4(0.f, 0.f, 0.f, 0.f)
                   : make_float4(D0 > 0.f ? T0 * expf(-D0) : T0, D1 > 0.f ? T1 * expf(-D1) : T1,
                                 D2 > 0.f ? T2 * expf(-D2) : T2, 0.f);
}

//! transmittance_walk over an instanced scene: the same passes, with the top level and the meshes walked in
//! one loop (scene_walk) and the candidates ordered by (t, instance, triangle); ra, rb = world ray
template<class F = alpha_filter>
__device__ float4 transmittance_walk_scene(const pt_tlas &tl, const gpu_material *__restrict__ mats, const float4 &ra,
                                           const float4 &rb, bool cull, int32_t *lds, const F &flt)
{
    int32_t spill[SCENE_STACK - PT_LDS];
    lane_stack<PT_BLOCK, PT_LDS> st{lds, spill};
    const ray_setup rw = make_ray_setup(ra, rb);
    float T0 = 1.f, T1 = 1.f, T2 = 1.f, D0 = 0.f, D1 = 0.f, D2 = 0.f;
    bool blocked = false, more = !tl.empty, first_pass = true;
    float lb_t = 0.f;
    uint64_t lb_id = 0;
    while(more && !blocked)
    {
        float kt[TK];
        uint64_t kid[TK];
        uint32_t km[TK];
#pragma unroll
        for(int j = 0; j < TK; ++j)
        {
            kt[j] = __int_as_float(0x7f800000);
            kid[j] = ~0ull;
            km[j] = 0u;
        }
        int cnt = 0;
        more = false;
        float bound = rw.tmax;
        const float lo = first_pass ? rw.tmin : fmaxf(rw.tmin, lb_t);
        ray_setup r = rw;
        const vkrt_bvh4_node *nodes = tl.tlas;
        const float4 *tris = nullptr;
        uint32_t inst = 0;
        int base = -1;
        int sp = 0;
        int32_t ref = tl.root;
        while(true)
        {
            if(ref >= 0)
            {
                int32_t c[4];
                const int nc = visit4(nodes, ref, r, bound, c, lo);
                if(nc)
                {
                    for(int j = nc - 1; j >= 1; --j) st.put(sp++, c[j]);
                    ref = c[0];
                    continue;
                }
            }
            else if(base < 0)
            {
                uint32_t first, count;
                leaf_decode(ref, first, count);
                inst = tl.order[first];
                const inst_dev I = tl.insts[inst];
                const blas_dev B = tl.blases[I.blas];
                if(!B.empty)
                {
                    r = object_ray(I, ra, rb);
                    nodes = B.nodes;
                    tris = B.tris;
                    base = sp;
                    ref = B.root;
                    continue;
                }
            }
            else
            {
                uint32_t first, count;
                leaf_decode(ref, first, count);
                for(uint32_t k = 0; k < count; ++k)
                {
                    const size_t tb = 48 * (size_t)(first + k);
                    const float4 p0 = load_global_f4(tris, tb), p1 = load_global_f4(tris, tb + 16),
                                 p2 = load_global_f4(tris, tb + 32);
                    const bool two_sided = (__float_as_uint(p2.w) & TRI_TWO_SIDED) != 0u;
                    float t, u, v;
                    if(!vierkant::rt::mt_fast(r.o, r.d, vec3{p0.x, p0.y, p0.z}, vec3{p0.w, p1.x, p1.y},
                                              vec3{p1.z, p1.w, p2.x}, t, u, v, cull && !two_sided) ||
                       !(t > r.tmin && t < r.tmax))
                        continue;
                    const uint32_t mid = __float_as_uint(p2.z);
                    const gpu_material m = mats[mid];
                    if(!(m.mrti.z > 0.0f))
                    {
                        if(filtered_out(flt, p2, u, v)) continue;
                        blocked = true;
                        break;
                    }
                    const uint64_t id = ((uint64_t)inst << 32) | __float_as_uint(p2.y);
                    const uint32_t tag = candidate_tag(mid, m, r.d, p0, p1, p2);
                    if(!first_pass && (t < lb_t || (t == lb_t && id <= lb_id))) continue;
                    if(cnt == TK && (t > kt[TK - 1] || (t == kt[TK - 1] && id > kid[TK - 1])))
                    {
                        more = true;
                        continue;
                    }
                    if(cnt == TK) more = true;
                    float ct = t;
                    uint64_t cid = id;
                    uint32_t cm = tag;
#pragma unroll
                    for(int j = 0; j < TK; ++j)
                    {
                        const bool less = ct < kt[j] || (ct == kt[j] && cid < kid[j]);
                        const float tt = kt[j];
                        const uint64_t ti = kid[j];
                        const uint32_t tm = km[j];
                        kt[j] = less ? ct : tt;
                        kid[j] = less ? cid : ti;
                        km[j] = less ? cm : tm;
                        ct = less ? tt : ct;
                        cid = less ? ti : cid;
                        cm = less ? tm : cm;
                    }
                    cnt = cnt < TK ? cnt + 1 : TK;
                    if(cnt == TK)
                    {
                        bound = kt[TK - 1];
                        more = true;
                    }
                }
                if(blocked) break;
            }
            if(sp == base)
            {
                r = rw;
                nodes = tl.tlas;
                base = -1;
            }
            if(sp == 0) break;
            --sp;
            ref = st.get(sp);
        }
        if(blocked) break;
#pragma unroll
        for(int j = 0; j < TK; ++j)
        {
            if(j < cnt)
            {
                const gpu_material m = mats[km[j] & ~KM_FRONT];
                T0 = T0 * (m.mrti.z * m.color.x);
                T1 = T1 * (m.mrti.z * m.color.y);
                T2 = T2 * (m.mrti.z * m.color.z);
                if(m.vol.w > 0.0f)
                {
                    const bool front = (km[j] & KM_FRONT) != 0u;
                    D0 += front ? -m.vol.x * kt[j] : m.vol.x * kt[j];
                    D1 += front ? -m.vol.y * kt[j] : m.vol.y * kt[j];
                    D2 += front ? -m.vol.z * kt[j] : m.vol.z * kt[j];
                }
                lb_t = kt[j];
                lb_id = kid[j];
            }
        }
        first_pass = false;
    }
    // unpaired entries (the light inside a volume) must not amplify: the depth is clamped at 0
    // (exp(-0) = 1 exactly: scenes without volumes skip the exp)
    return blocked ? make_float4(0.f, 0.f, 0.f, 0.f)
                   : make_float4(D0 > 0.f ? T0 * expf(-D0) : T0, D1 > 0.f ? T1 * expf(-D1) : T1,
                                 D2 > 0.f ? T2 * expf(-D2) : T2, 0.f);
}

template<bool INST>
__global__ void __launch_bounds__(PT_BLOCK) k_pt_transmittance(const vkrt_bvh_node *__restrict__ nodes2,
                                                               int32_t root2, const vkrt_bvh4_node *__restrict__ nodes4,
                                                               int32_t root4, bool wide, bool empty,
                                                               const float4 *__restrict__ tris,
                                                               const gpu_material *__restrict__ mats,
                                                               const float4 *__restrict__ rays, uint32_t n, bool cull,
                                                               float4 *__restrict__ out, alpha_filter flt, pt_tlas tl)
{
    __shared__ int32_t smem[PT_LDS * PT_BLOCK];
    const uint32_t i = blockIdx.x * PT_BLOCK + threadIdx.x;
    if(i >= n) return;
    if constexpr(INST)
        out[i] = transmittance_walk_scene(tl, mats, rays[2 * (size_t)i], rays[2 * (size_t)i + 1], cull,
                                          smem + threadIdx.x, flt);
    else
        out[i] = transmittance_walk(nodes2, root2, nodes4, root4, wide, empty, tris, mats, setup_ray(rays, i), cull,
                                    smem + threadIdx.x, flt);
}

//! trace one shadow record and add its next-event estimate to the pixel (and to its training vertex)
template<class F = alpha_filter, bool INST = false>
__device__ __forceinline__ void shadow_path(const vkrt_bvh_node *__restrict__ nodes2, int32_t root2,
                                            const vkrt_bvh4_node *__restrict__ nodes4, int32_t root4, bool wide,
                                            bool empty, const float4 *__restrict__ tris,
                                            const gpu_material *__restrict__ mats, const pt_frame &f,
                                            const pt_buffers &b, const float4 (&rec)[SHADOW_F4], int32_t *smem,
                                            F flt, const pt_tlas &tl = pt_tlas{})
{
    const float4 o = rec[0], d = rec[1];
    if constexpr(F::active)
    {
        flt.per_candidate = true;
        flt.seed = __float_as_uint(rec[5].z);
    }
    float4 T;
    if constexpr(INST) T = transmittance_walk_scene(tl, mats, o, d, false, smem, flt);
    else T = transmittance_walk(nodes2, root2, nodes4, root4, wide, empty, tris, mats, make_ray_setup(o, d), false, smem,
                                flt);
    if(!(T.x > 0.0f || T.y > 0.0f || T.z > 0.0f)) return;
    const float4 lr = rec[2], fc = rec[3], bp = rec[4], id = rec[5];
    vec3 nee = vec3{T.x, T.y, T.z} * lr.w;
    nee = nee * static_cast<float>(f.n_lights);
    nee = mul3(nee, xyz(lr));
    nee = mul3(nee, xyz(fc));
    nee = nee * fc.w;
    nee = vec3{nee.x / bp.w, nee.y / bp.w, nee.z / bp.w};
    const vec3 c = mul3(xyz(bp), nee);
    const uint32_t p = __float_as_uint(id.x), vtx = __float_as_uint(id.y);
    float4 a = b.acc[p];
    b.acc[p] = make_float4(a.x + c.x, a.y + c.y, a.z + c.z, a.w);
    if(vtx != NO_VERTEX)
    {
        float4 *e = b.tvtx + (size_t)vtx * VERTEX_F4 + 4;
        const float4 ev = *e;
        *e = make_float4(ev.x + nee.x, ev.y + nee.y, ev.z + nee.z, ev.w);
    }
}

//! the shadow half of nee_surface (direct_lighting.slang:311-333): the unculled transmittance toward the
//! sampled light, then transmittance * weight * n_lights * radiance * F * cos / pdf scaled by the vertex's
//! throughput joins the pixel's sample sum; a training vertex also records the unscaled estimate
template<class F, bool INST>
__global__ void __launch_bounds__(PT_BLOCK)
__attribute__((amdgpu_waves_per_eu(INST || F::active ? 1 : PT_SHADOW_PLAIN_WAVES_PER_SIMD))) k_pt_shadow(const vkrt_bvh_node *__restrict__ nodes2, int32_t root2,
                                                        const vkrt_bvh4_node *__restrict__ nodes4, int32_t root4,
                                                        bool wide, bool empty, const float4 *__restrict__ tris,
                                                        const gpu_material *__restrict__ mats, pt_frame f,
                                                        pt_buffers b, F flt, pt_tlas tl)
{
    __shared__ int32_t smem[PT_LDS * PT_BLOCK];
    const uint32_t count = b.counter[6];
    for(uint32_t i = blockIdx.x * PT_BLOCK + threadIdx.x; i < count; i += gridDim.x * PT_BLOCK)
    {
        const float4 *rec = b.shadow + (size_t)i * SHADOW_F4;
        float4 r[SHADOW_F4];
        for(uint32_t k = 0; k < SHADOW_F4; ++k) r[k] = rec[k];
        shadow_path<F, INST>(nodes2, root2, nodes4, root4, wide, empty, tris, mats, f, b, r, smem + threadIdx.x, flt,
                             tl);
    }
}

//! the cache mode's training paths, one lane per path through every depth: trace, shade and the shadow ray
//! of the next-event estimate, with no kernel boundary between depths. Launched on the second stream at the
//! start of a sample, its few waves stay resident beside the main paths' full-width launches, so the chain of
//! dependent traversals does not wait for free compute units at each depth
template<bool LIGHTS, bool INST>
__global__ void __launch_bounds__(PT_BLOCK) k_pt_train_paths(const vkrt_bvh_node *__restrict__ nodes2,
                                                             const float4 *__restrict__ tris, int32_t root2,
                                                             bool empty, const vkrt_bvh4_node *__restrict__ nodes4,
                                                             int32_t root4, bool wide, pt_frame f, pt_buffers b,
                                                             pt_scene_dev s, pt_tlas tl)
{
    __shared__ int32_t smem[PT_LDS * PT_BLOCK];
    const uint32_t q = blockIdx.x * PT_BLOCK + threadIdx.x;
    if(q >= b.counter[0]) return;
    const uint32_t p = b.queue[0][q];
    const alpha_filter flt{s.idx, s.uv, s.mats, s.tex, s.texels, f.disable_material != 0u};
    const alpha_filter shadow_flt{s.idx, s.uv, s.mats, s.tex, s.texels, false};
    uint32_t segments = 0;
    for(uint32_t depth = 0; depth < f.max_depth; ++depth)
    {
        trace_path<PT_BLOCK, alpha_filter, INST>(nodes2, tris, root2, empty, nodes4, root4, wide, b, p,
                                                 smem + threadIdx.x, flt, depth, tl);
        bool alive = false, to_cache = false, need_shadow = false;
        float4 shadow_rec[SHADOW_F4];
        float4 qfeat[QUERY_F4];
        shade_path<LIGHTS, INST>(f, b, s, depth, p, alive, to_cache, need_shadow, shadow_rec, qfeat);
        ++segments;
        if(LIGHTS && need_shadow)
            shadow_path<alpha_filter, INST>(nodes2, root2, nodes4, root4, wide, empty, tris, s.mats, f, b, shadow_rec,
                                            smem + threadIdx.x, shadow_flt, tl);
        if(!alive) break;
    }
    atomicAdd(reinterpret_cast<unsigned long long *>(&b.counter[2]), (unsigned long long)segments);
}

//! the deep-bounce tail: the paths alive at depth d0 (queue `in`) finish in one lane each, trace, shade and
//! shadow ray at every remaining depth with no kernel boundary between depths. Past the first few depths the
//! live queue fills a fraction of the GPU and each per-depth launch lasts as long as its slowest wave's
//! traversal (a chain of dependent node loads); here a wave's lanes run their depths back to back and the
//! waves overlap. Each path performs the operations of the per-depth kernels in the same order (trace, shade,
//! shadow ray, next depth), so the image is bit for bit the same.
template<bool LIGHTS, class F, bool INST>
__global__ void __launch_bounds__(PT_BLOCK) k_pt_finish(const vkrt_bvh_node *__restrict__ nodes2,
                                                        const float4 *__restrict__ tris, int32_t root2, bool empty,
                                                        const vkrt_bvh4_node *__restrict__ nodes4, int32_t root4,
                                                        bool wide, pt_frame f, pt_buffers b, pt_scene_dev s, F flt,
                                                        F shadow_flt, uint32_t d0, int in, pt_tlas tl)
{
    __shared__ int32_t smem[PT_LDS * PT_BLOCK];
    __shared__ unsigned long long s_segments;
    if(threadIdx.x == 0) s_segments = 0ull;
    __syncthreads();
    const uint32_t q = blockIdx.x * PT_BLOCK + threadIdx.x;
    const uint32_t count = b.counter[in];
    uint32_t segments = 0;
    if(q < count)
    {
        const uint32_t p = b.queue[in][q];
        for(uint32_t depth = d0; depth < f.max_depth; ++depth)
        {
            trace_path<PT_BLOCK, F, INST>(nodes2, tris, root2, empty, nodes4, root4, wide, b, p, smem + threadIdx.x,
                                          flt, depth, tl);
            bool alive = false, to_cache = false, need_shadow = false;
            float4 shadow_rec[SHADOW_F4];
            float4 qfeat[QUERY_F4];
            shade_path<LIGHTS, INST>(f, b, s, depth, p, alive, to_cache, need_shadow, shadow_rec, qfeat);
            ++segments;
            if(LIGHTS && need_shadow)
                shadow_path<F, INST>(nodes2, root2, nodes4, root4, wide, empty, tris, s.mats, f, b, shadow_rec,
                                     smem + threadIdx.x, shadow_flt, tl);
            if(!alive) break;
        }
    }
    if(segments) atomicAdd(&s_segments, (unsigned long long)segments);
    __syncthreads();
    if(threadIdx.x == 0 && s_segments)
        atomicAdd(reinterpret_cast<unsigned long long *>(&b.counter[2]), s_segments);
}

// ---------------------------------------------------------------------------------------------
// denoise pass (shaders/slang/denoise.slang, PBRPathTracer::denoise_pass src/PBRPathTracer.cpp:346-400):
// edge-avoiding a-trous wavelet over the pixel_data_t buffer, steps 1, 2, 4, 8, 16 ping-ponging between two
// float4 images; without denoising the image is the raw radiance and alpha
// ---------------------------------------------------------------------------------------------
constexpr uint32_t ATROUS_STEPS[5] = {1, 2, 4, 8, 16};

//! exposure/gamma parameters of the tonemap stage; `on` is 0 when the image stays linear
struct tonemap_params
{
    int on;
    float exposure, inv_gamma;
};


__device__ inline float4 tonemap_texel(float4 c, const tonemap_params &tm)
{
    if(!tm.on) return c;
    return make_float4(half_value(tonemap_channel(c.x, tm.exposure, tm.inv_gamma)),
                       half_value(tonemap_channel(c.y, tm.exposure, tm.inv_gamma)),
                       half_value(tonemap_channel(c.z, tm.exposure, tm.inv_gamma)), half_value(c.w));
}

//! one pass of denoise.slang (16 x 16 pixels per workgroup, as its numthreads); the last pass of a chain
//! also applies the tonemap stage, so the composited image costs no extra read and write
__global__ void __launch_bounds__(256) k_denoise(const vkrt_pixel *__restrict__ px, const float4 *__restrict__ prev,
                                                 float4 *__restrict__ out, uint32_t w, uint32_t h, int step,
                                                 int denoise, tonemap_params tm)
{
    const uint32_t x = blockIdx.x * 16 + threadIdx.x % 16, y = blockIdx.y * 16 + threadIdx.x / 16;
    if(x >= w || y >= h) return;
    const uint32_t i = y * w + x;
    if(!denoise)
    {
        out[i] = tonemap_texel(make_float4(px[i].radiance[0], px[i].radiance[1], px[i].radiance[2], px[i].alpha),
                               tm);
        return;
    }
    const vec3 c = atrous_pixel(px, reinterpret_cast<const float *>(prev), w, h, static_cast<int>(x),
                                static_cast<int>(y), step);
    out[i] = tonemap_texel(make_float4(c.x, c.y, c.z, 0.f), tm);
}

__global__ void __launch_bounds__(256) k_tonemap(const float4 *in, float4 *out, size_t n,
                                                 tonemap_params tm)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if(i < n) out[i] = tonemap_texel(in[i], tm);
}

static tonemap_params make_tonemap(bool on, float exposure, float gamma)
{
    VKRT_REQUIRE(!on || (gamma > 0.f && std::isfinite(gamma)), "gamma must be positive and finite");
    VKRT_REQUIRE(!on || std::isfinite(exposure), "exposure must be finite");
    // bloom_composition.slang:44 raises to float3(1.0 / gamma), a single-precision quotient
    return tonemap_params{on ? 1 : 0, exposure, on ? 1.0f / gamma : 1.0f};
}

//! the image of a pixel buffer: five a-trous passes when denoising, the last one tonemapping when asked;
//! returns the buffer holding the result
static float4 *run_denoise(hipStream_t s, const vkrt_pixel *px, uint32_t w, uint32_t h, bool denoise,
                           const tonemap_params &tm, float4 *ping, float4 *pong)
{
    const dim3 grid(div_up(w, 16), div_up(h, 16));
    float4 *bufs[2] = {ping, pong};
    const uint32_t passes = denoise ? 5u : 1u;
    const tonemap_params off{0, 0.f, 1.f};
    for(uint32_t p = 0; p < passes; ++p)
        k_denoise<<<grid, 256, 0, s>>>(px, p ? bufs[(p - 1) % 2] : nullptr, bufs[p % 2], w, h,
                                       static_cast<int>(ATROUS_STEPS[p]), denoise ? 1 : 0,
                                       p + 1 == passes ? tm : off);
    VKRT_HIP(hipGetLastError());
    return bufs[(passes - 1) % 2];
}

//! the leaf records' material and flags; prim_offset turns a mesh's own triangle numbers into the scene's
//! (an instanced scene's meshes are triangle ranges of one list)
__global__ void k_pt_tag(float4 *__restrict__ tris, uint32_t n, const uint32_t *__restrict__ tri_object,
                         const uint32_t *__restrict__ object_material, const gpu_material *__restrict__ mats,
                         uint32_t prim_offset)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if(j >= n) return;
    float4 r = tris[3 * (size_t)j + 2];
    const uint32_t prim = __float_as_uint(r.y) + prim_offset;
    r.y = __uint_as_float(prim);
    const uint32_t mat = object_material[tri_object[prim]];
    r.z = __uint_as_float(mat);
    r.w = __uint_as_float((mats[mat].sheen.w != 0.f ? TRI_TWO_SIDED : 0u) |
                          (mats[mat].tex.y != VKRT_ALPHA_OPAQUE ? TRI_ALPHA_MASK : 0u));
    tris[3 * (size_t)j + 2] = r;
}

}// namespace vkrt

// =============================================================================================
// the tracer object
// =============================================================================================
struct vkrt_pt
{
    vkrt_ctx *ctx = nullptr;
    vkrt_pt_settings settings{};
    int launch_mode = VKRT_PT_LAUNCH_FULL_GRID;// per-depth launch shape (vkrt_pt_set_launch_mode)
    uint32_t tail_depth = vkrt::PT_TAIL_DEPTH;  // survivors at this depth finish in one lane (k_pt_finish)
    bool any_alpha = false;                    // some material is alpha-tested (MASK or BLEND)
    bool any_volume = false;                   // some material has an interior medium
    // instanced scenes (vkrt_pt_set_instanced_scene): one BVH per mesh, the top level over the instances, per
    // instance its shading transform, per object its object-space half extents; `bounds` is the scene's world
    // box either way (the cache grid's extent)
    bool instanced = false;
    std::vector<vkrt_bvh *> mesh_bvhs;
    vkrt_scene *scene = nullptr;
    vkrt::pt_tlas tlas{};
    vkrt::device_buffer inst_xf, object_he, hit_inst;
    float bounds[6] = {0.f, 0.f, 0.f, -1.f, -1.f, -1.f};
    std::minstd_rand0 random_engine;// std::default_random_engine of libstdc++ (PBRPathTracer.cpp:32)
    uint32_t batch_index = 0;
    vkrt_bvh *bvh = nullptr;
    uint32_t n_tri = 0, n_vert = 0, n_objects = 0, n_materials = 0;
    float environment = 1.f;
    vkrt::device_buffer pos, nrm, idx, tri_object, object_material, object_tmax, mats;
    // texture coordinates, texture descriptors and their texels (all mip levels)
    vkrt::device_buffer uv, tan, tex_desc_buf, texels;
    uint32_t width = 0, height = 0;
    vkrt::device_buffer ray_o, ray_d, hit, beta, rad, state, iors, media, acc, info, queue0, queue1, counter, pixels, cone;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the cache mode's training paths: their own queues, traced on a second stream beside the main paths
    hipStream_t train_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    vkrt::device_buffer train_queue0, train_queue1;
    vkrt_pt_stats stats{};
    // radiance cache (not owned) and its per-sample buffers
    vkrt_rc *rc = nullptr;
    vkrt_pt_cache_config cache_cfg{};
    vkrt::device_buffer qrec, tvtx, trec, ttgt;
    // analytic lights (the frame's light list) and the shadow records of one depth
    vkrt::device_buffer lights, shadow;
    uint32_t n_lights = 0;
    vkrt::device_buffer image;// denoise ping-pong images
};

namespace vkrt
{

static void default_settings(vkrt_pt_settings &s)
{
    s = {};
    s.resolution[0] = 1280;
    s.resolution[1] = 720;
    s.max_num_batches = 0;
    s.num_samples = 1;
    s.max_trace_depth = 6;
    s.max_path_beta = 0.f;
    s.suspend_trace_when_done = 1;
    s.disable_material = 0;
    s.suppress_reset = 0;
    s.denoising = 0;
    s.tonemap = 1;// settings_t defaults (PBRPathTracer.hpp:77-90)
    s.gamma = 1.0f;
    s.exposure = 2.0f;
}

static void check_settings(const vkrt_pt_settings &s)
{
    VKRT_REQUIRE(s.resolution[0] > 0 && s.resolution[1] > 0, "resolution must be positive");
    VKRT_REQUIRE((uint64_t)s.resolution[0] * s.resolution[1] <= (1u << 26), "at most 2^26 pixels");
    VKRT_REQUIRE(s.num_samples > 0 && s.num_samples <= 65536, "num_samples must be in 1..65536");
    VKRT_REQUIRE(s.max_trace_depth > 0 && s.max_trace_depth <= 64, "max_trace_depth must be in 1..64");
    VKRT_REQUIRE(!s.tonemap || (s.gamma > 0.f && std::isfinite(s.gamma) && std::isfinite(s.exposure)),
                 "tonemap needs a positive finite gamma and a finite exposure");
    VKRT_REQUIRE(s.mis_mode <= VKRT_MIS_BSDF_ONLY, "mis_mode must be 0, 1 or 2");
}

//! resize_storage (PBRPathTracer.cpp:684-830): new pixel storage resets the batch index
static void resize_storage(vkrt_pt *pt, uint32_t w, uint32_t h)
{
    if(w == pt->width && h == pt->height) return;
    const size_t n = (size_t)w * h;
    pt->ray_o.reserve(n * 16);
    pt->ray_d.reserve(n * 16);
    pt->hit.reserve(n * 16);
    pt->hit_inst.reserve(n * 4);
    pt->beta.reserve(n * 16);
    pt->rad.reserve(n * 16);
    pt->state.reserve(n * 16);
    pt->iors.reserve(n * 16);
    pt->media.reserve(n * 8);
    pt->cone.reserve(n * 8);
    pt->acc.reserve(n * 16);
    pt->info.reserve(n * 16);
    pt->queue0.reserve(n * 4);
    pt->queue1.reserve(n * 4);
    pt->counter.reserve(PT_COUNTER_WORDS * sizeof(uint32_t));
    pt->pixels.reserve(n * sizeof(vkrt_pixel));
    VKRT_HIP(hipMemsetAsync(pt->pixels.ptr, 0, n * sizeof(vkrt_pixel), pt->ctx->stream));
    pt->width = w;
    pt->height = h;
    pt->batch_index = 0;
}

static pt_buffers buffers(vkrt_pt *pt)
{
    pt_buffers b;
    b.ray_o = pt->ray_o.as<float4>();
    b.ray_d = pt->ray_d.as<float4>();
    b.hit = pt->hit.as<uint4>();
    b.beta = pt->beta.as<float4>();
    b.rad = pt->rad.as<float4>();
    b.state = pt->state.as<uint4>();
    b.iors = pt->iors.as<float4>();
    b.media = pt->media.as<uint2>();
    b.cone = pt->cone.as<float2>();
    b.acc = pt->acc.as<float4>();
    b.info = pt->info.as<uint4>();
    b.queue[0] = pt->queue0.as<uint32_t>();
    b.queue[1] = pt->queue1.as<uint32_t>();
    b.counter = pt->counter.as<uint32_t>();
    b.pixels = pt->pixels.as<vkrt_pixel>();
    b.qrec = b.tvtx = b.trec = b.ttgt = b.shadow = nullptr;
    b.train_queue = nullptr;
    b.hit_inst = pt->hit_inst.as<uint32_t>();
    return b;
}

//! the scene's device arrays as the kernels take them
static pt_scene_dev scene_dev(vkrt_pt *pt)
{
    return pt_scene_dev{pt->pos.as<float>(), pt->nrm.as<float>(), pt->idx.as<uint32_t>(),
                        pt->tri_object.as<uint32_t>(), pt->object_material.as<uint32_t>(), pt->object_tmax.as<float>(),
                        pt->mats.as<gpu_material>(), pt->lights.as<vkrt_light>(), pt->uv.as<float2>(),
                        pt->tex_desc_buf.as<tex_desc>(), pt->texels.as<uint32_t>(), pt->tan.as<float>(),
                        pt->inst_xf.as<float4>(), pt->tlas.insts, pt->object_he.as<float>()};
}

//! f32tof16 on the host: round to nearest even at the binade's spacing of halves (2^-24 below 2^-14),
//! overflow to infinity
static float half_round_host(float x)
{
    if(!std::isfinite(x) || x == 0.f) return x;
    int k;
    std::frexp(x, &k);
    const int q = (k - 1 >= -14) ? (k - 1) - 10 : -24;
    const double r = std::ldexp(std::nearbyint(std::ldexp((double)x, -q)), q);
    if(std::fabs(r) > 65504.0) return std::copysign(INFINITY, x);
    return static_cast<float>(r);
}

static pt_frame make_frame(vkrt_pt *pt, const vkrt_camera &cam, uint32_t seed)
{
    pt_frame f;
    f.width = pt->width;
    f.height = pt->height;
    f.n_pix = pt->width * pt->height;
    f.num_samples = pt->settings.num_samples;
    f.max_depth = pt->settings.max_trace_depth;
    f.seed = seed;
    f.batch_index = pt->batch_index;
    f.disable_material = pt->settings.disable_material;
    f.max_path_beta = pt->settings.max_path_beta;
    f.environment = pt->environment;
    // camera basis in double precision, rounded once to float: forward, right = forward x up, up = right x forward
    auto nrm = [](const double v[3], double out[3]) {
        const double l = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        for(int k = 0; k < 3; ++k) out[k] = v[k] / l;
    };
    auto crs = [](const double a[3], const double b[3], double out[3]) {
        out[0] = a[1] * b[2] - a[2] * b[1];
        out[1] = a[2] * b[0] - a[0] * b[2];
        out[2] = a[0] * b[1] - a[1] * b[0];
    };
    const double fw0[3] = {cam.forward[0], cam.forward[1], cam.forward[2]};
    const double up0[3] = {cam.up[0], cam.up[1], cam.up[2]};
    double fw[3], rt0[3], rt[3], up[3];
    nrm(fw0, fw);
    crs(fw, up0, rt0);
    nrm(rt0, rt);
    crs(rt, fw, up);
    f.eye = vec3{cam.eye[0], cam.eye[1], cam.eye[2]};
    f.forward = vec3{(float)fw[0], (float)fw[1], (float)fw[2]};
    f.right = vec3{(float)rt[0], (float)rt[1], (float)rt[2]};
    f.up = vec3{(float)up[0], (float)up[1], (float)up[2]};
    const double th = std::tan(0.5 * (double)cam.fov_y);
    f.tan_half_y = (float)th;
    f.tan_half_x = (float)(th * (double)pt->width / (double)pt->height);
    f.aperture = cam.aperture;
    f.focal_distance = cam.focal_distance > 0.f ? cam.focal_distance : 1.f;
    // the ray cone's spread angle (raypipeline.slang:102), kept as a half like the payload's packed cone
    f.cone_spread = half_round_host(std::atan(2.0f * std::tan(cam.fov_y / 2.0f) / static_cast<float>(pt->height)));
    f.cache_on = pt->rc ? 1u : 0u;
    f.cache_depth = pt->cache_cfg.cache_depth;
    f.train_stride = std::max<uint32_t>(1u, pt->cache_cfg.train_stride);
    f.train_offset = 0;
    f.n_train_paths = div_up(f.n_pix, f.train_stride);
    f.train_max = pt->cache_cfg.max_train_records;
    f.n_lights = pt->n_lights;
    f.mis_mode = pt->settings.mis_mode;
    f.suppress_caustics = pt->settings.suppress_refractive_caustics ? 1u : 0u;
    f.any_volume = pt->any_volume ? 1u : 0u;
    return f;
}

//! the cache's grid covers the scene bounds, padded by 1% per side
static void fit_cache_bounds(vkrt_pt *pt)
{
    if(!pt->rc || !(pt->bvh || pt->instanced)) return;
    const float *bb = pt->bounds;
    float lo[3], hi[3];
    for(int a = 0; a < 3; ++a)
    {
        const bool valid = bb[a + 3] >= bb[a];
        const float e = valid ? bb[a + 3] - bb[a] : 0.f;
        const float pad = std::max(0.01f * e, 1e-3f);
        lo[a] = valid ? bb[a] - pad : -1.f;
        hi[a] = valid ? bb[a + 3] + pad : 1.f;
    }
    const int rc = vkrt_rc_set_bounds(pt->rc, lo, hi);
    if(rc != VKRT_OK) throw error(rc, "radiance cache bounds: " + pt->ctx->last_error);
}

//! the filter object of instance F: the alpha filter, or keep_all
template<class F>
static F filter_of(const alpha_filter &a)
{
    if constexpr(std::is_same_v<F, alpha_filter>) return a;
    else return F{};
}

static void trace_batch(vkrt_pt *pt, const vkrt_camera &cam, uint32_t seed)
{
    hipStream_t s = pt->ctx->stream;
    pt_frame f = make_frame(pt, cam, seed);
    pt_buffers b = buffers(pt);
    const pt_scene_dev sc = scene_dev(pt);
    const alpha_filter flt{sc.idx, sc.uv, sc.mats, sc.tex, sc.texels, f.disable_material != 0u};
    const alpha_filter shadow_flt{sc.idx, sc.uv, sc.mats, sc.tex, sc.texels, false};
    const uint32_t blocks = div_up(f.n_pix, PT_BLOCK);
    const vkrt_bvh *bvh = pt->bvh;
    if(f.cache_on)
    {
        pt->qrec.reserve((size_t)f.n_pix * QUERY_F4 * 16);
        pt->tvtx.reserve((size_t)f.n_train_paths * f.max_depth * VERTEX_F4 * 16);
        pt->trec.reserve((size_t)std::max(1u, f.train_max) * 64);
        pt->ttgt.reserve((size_t)std::max(1u, f.train_max) * 16);
        b.qrec = pt->qrec.as<float4>();
        b.tvtx = pt->tvtx.as<float4>();
        b.trec = pt->trec.as<float4>();
        b.ttgt = pt->ttgt.as<float4>();
    }
    if(f.n_lights && !f.cache_on)
    {
        pt->shadow.reserve((size_t)f.n_pix * SHADOW_F4 * 16);
        b.shadow = pt->shadow.as<float4>();
    }
    // the training stream's view: its own queues and counters, its own half of the shadow records
    pt_buffers bt = b;
    if(f.cache_on)
    {
        const size_t n_train_max = div_up(f.n_pix, f.train_stride);
        pt->train_queue0.reserve(n_train_max * 4);
        pt->train_queue1.reserve(n_train_max * 4);
        b.train_queue = pt->train_queue0.as<uint32_t>();
        bt.queue[0] = pt->train_queue0.as<uint32_t>();
        bt.queue[1] = pt->train_queue1.as<uint32_t>();
        bt.counter = b.counter + TRAIN_COUNTERS;
        if(f.n_lights)
        {
            pt->shadow.reserve(((size_t)f.n_pix + n_train_max) * SHADOW_F4 * 16);
            b.shadow = pt->shadow.as<float4>();
            bt.shadow = b.shadow + (size_t)f.n_pix * SHADOW_F4;
        }
    }
    // the flat scene's BVH (an instanced scene has none: its kernels walk pt->tlas)
    const vkrt_bvh_node *nodes2 = bvh ? bvh->nodes.as<vkrt_bvh_node>() : nullptr;
    const vkrt_bvh4_node *nodes4 = bvh ? bvh->nodes4.as<vkrt_bvh4_node>() : nullptr;
    const float4 *tris = bvh ? bvh->tris.as<float4>() : nullptr;
    const int32_t root_ref = bvh ? bvh->root_ref : 0, root4 = bvh ? bvh->root4 : 0;
    const bool flat_empty = bvh ? bvh->empty : true, wide = bvh ? bvh->wide : false;
    // one depth of a queue of paths: trace, shade, and the shadow rays of next-event estimation
    // scenes with no alpha-tested material take the trace and shadow kernels without the any-hit filter
    const bool alpha = pt->any_alpha;
    // persistent grids of the trace and shadow kernels: the workgroups resident at once (the runtime's
    // occupancy figure for the kernel's registers and LDS), at most one per tile of a full queue
    // (VKRT_PT_LAUNCH_FULL_GRID launches one workgroup per tile of every pixel at every depth instead: the same
    // kernels, each workgroup then taking at most one tile)
    const bool full = pt->launch_mode == VKRT_PT_LAUNCH_FULL_GRID;
    const uint32_t cus = static_cast<uint32_t>(pt->ctx->cu_count);
    auto resident = [&](auto kernel, int block, uint32_t cap) {
        int per_cu = 0;
        VKRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
        return full ? cap : std::min<uint32_t>(cus * static_cast<uint32_t>(std::max(per_cu, 1)), cap);
    };
    const uint32_t trace_full = div_up(f.n_pix, PT_TRACE_BLOCK);
    // compile-time variants: any-hit filter or none, flat or instanced scene
    const bool inst = pt->instanced;
    const pt_tlas tl = pt->tlas;
    auto variant = [&](auto fn) {
        auto by_inst = [&](auto a) {
            if(inst) fn(a, std::true_type{});
            else fn(a, std::false_type{});
        };
        if(alpha) by_inst(std::true_type{});
        else by_inst(std::false_type{});
    };
    uint32_t trace_grid = 0, shadow_grid = 0;
    variant([&](auto a, auto in_) {
        using F = std::conditional_t<decltype(a)::value, alpha_filter, keep_all>;
        trace_grid = resident(k_pt_trace<F, decltype(in_)::value>, PT_TRACE_BLOCK, trace_full);
        shadow_grid = resident(k_pt_shadow<F, decltype(in_)::value>, PT_BLOCK, blocks);
    });
    auto depth_step = [&](hipStream_t st, const pt_buffers &bb, uint32_t depth, int in) {
        k_pt_reset_depth<<<1, XCDS, 0, st>>>(bb.counter, in ^ 1);
        variant([&](auto a, auto in_) {
            using F = std::conditional_t<decltype(a)::value, alpha_filter, keep_all>;
            constexpr bool I = decltype(in_)::value;
            k_pt_trace<F, I><<<trace_grid, PT_TRACE_BLOCK, 0, st>>>(nodes2, tris, root_ref, flat_empty, nodes4,
                                                                    root4, wide, bb, in, filter_of<F>(flt),
                                                                    depth, tl);
            if(f.n_lights)
            {
                k_pt_shade<true, I><<<blocks, PT_BLOCK, 0, st>>>(f, bb, sc, depth, in);
                k_pt_shadow<F, I><<<shadow_grid, PT_BLOCK, 0, st>>>(nodes2, root_ref, nodes4, root4,
                                                                    wide, flat_empty, tris, sc.mats, f, bb,
                                                                    filter_of<F>(shadow_flt), tl);
            }
            else
                k_pt_shade<false, I><<<blocks, PT_BLOCK, 0, st>>>(f, bb, sc, depth, in);
        });
    };
    VKRT_HIP(hipEventRecord(pt->ev0, s));
    k_pt_zero<<<1, 2 * TRAIN_COUNTERS, 0, s>>>(b.counter);
    for(uint32_t sample = 0; sample < f.num_samples; ++sample)
    {
        uint32_t main_depths = f.max_depth;
        if(f.cache_on)
        {
            // training pixels move every sample: offset from the batch seed and the sample index
            f.train_offset = xxhash32(seed, sample) % f.train_stride;
            f.n_train_paths = f.train_offset < f.n_pix ? div_up(f.n_pix - f.train_offset, f.train_stride) : 0u;
            // paths other than training paths end at the cache depth
            main_depths = std::min(f.max_depth, f.cache_depth + 1u);
        }
        k_pt_begin<<<blocks, PT_BLOCK, 0, s>>>(f, b, sample);
        const bool fork = f.cache_on && f.n_train_paths;
        // the deep-bounce tail (not in the cache mode, whose main paths end at the cache depth and queue cache
        // queries): the survivors at tail_depth finish in one lane each
        const uint32_t per_depth = f.cache_on ? main_depths : std::min(main_depths, pt->tail_depth);
        if(fork) VKRT_HIP(hipEventRecord(pt->ev_fork, s));
        // the first depth is queued before the training stream's launches, so the main stream does not wait
        // for the host to queue them
        if(per_depth > 0) depth_step(s, b, 0, 0);
        if(fork)
        {
            // the training paths' long, narrow tail is latency-bound: it runs on the second stream while the
            // main stream traces the full-width depths, followed there by the training step (targets, gradients,
            // Adam), which no main-path kernel reads before the cache query; the query waits for it, so the
            // main paths terminate into the cache as trained on this sample's training paths
            VKRT_HIP(hipStreamWaitEvent(pt->train_stream, pt->ev_fork, 0));
            const uint32_t tgrid = div_up(f.n_train_paths, PT_BLOCK);
            auto train = [&](auto lights, auto in_) {
                k_pt_train_paths<decltype(lights)::value, decltype(in_)::value>
                        <<<tgrid, PT_BLOCK, 0, pt->train_stream>>>(nodes2, tris, root_ref, flat_empty, nodes4,
                                                                   root4, wide, f, bt, sc, tl);
            };
            if(f.n_lights && inst) train(std::true_type{}, std::true_type{});
            else if(f.n_lights) train(std::true_type{}, std::false_type{});
            else if(inst) train(std::false_type{}, std::true_type{});
            else train(std::false_type{}, std::false_type{});
            if(f.train_max)
            {
                k_pt_train_targets<<<tgrid, PT_BLOCK, 0, pt->train_stream>>>(f, b);
                rc::train(pt->rc, b.trec, 4, b.ttgt, f.train_max, false, b.counter + 5, pt->train_stream);
            }
            VKRT_HIP(hipGetLastError());
            VKRT_HIP(hipEventRecord(pt->ev_join, pt->train_stream));
        }
        for(uint32_t depth = 1; depth < per_depth; ++depth) depth_step(s, b, depth, static_cast<int>(depth & 1u));
        if(per_depth < main_depths && per_depth > 0)
        {
            const int in = static_cast<int>(per_depth & 1u);
            variant([&](auto a, auto in_) {
                using F = std::conditional_t<decltype(a)::value, alpha_filter, keep_all>;
                constexpr bool I = decltype(in_)::value;
                if(f.n_lights)
                    k_pt_finish<true, F, I><<<blocks, PT_BLOCK, 0, s>>>(
                        nodes2, tris, root_ref, flat_empty, nodes4, root4, wide, f, b, sc,
                        filter_of<F>(flt), filter_of<F>(shadow_flt), per_depth, in, tl);
                else
                    k_pt_finish<false, F, I><<<blocks, PT_BLOCK, 0, s>>>(
                        nodes2, tris, root_ref, flat_empty, nodes4, root4, wide, f, b, sc,
                        filter_of<F>(flt), filter_of<F>(shadow_flt), per_depth, in, tl);
            });
        }
        if(f.cache_on)
        {
            if(fork) VKRT_HIP(hipStreamWaitEvent(s, pt->ev_join, 0));
            rc::query(pt->rc, b.qrec, QUERY_F4, f.n_pix, nullptr, b.acc, b.counter + 4);
        }
    }
    k_pt_resolve<<<blocks, PT_BLOCK, 0, s>>>(f, b);
    VKRT_HIP(hipEventRecord(pt->ev1, s));
    VKRT_HIP(hipGetLastError());
    uint64_t seg[2] = {0, 0};
    VKRT_HIP(hipMemcpyAsync(&seg[0], b.counter + 2, 8, hipMemcpyDeviceToHost, s));
    VKRT_HIP(hipMemcpyAsync(&seg[1], b.counter + TRAIN_COUNTERS + 2, 8, hipMemcpyDeviceToHost, s));
    VKRT_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    VKRT_HIP(hipEventElapsedTime(&ms, pt->ev0, pt->ev1));
    pt->stats.raytrace_ms = ms;
    pt->stats.traced = 1;
    pt->stats.segments = seg[0] + seg[1];
    pt->stats.samples = (uint64_t)f.n_pix * f.num_samples;
}

}// namespace vkrt

using namespace vkrt;

extern "C" {

void vkrt_pt_default_settings(vkrt_pt_settings *out)
{
    if(out) default_settings(*out);
}

int vkrt_pt_create(vkrt_ctx *ctx, const vkrt_pt_settings *settings, uint32_t seed, vkrt_pt **out)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(out, "null output handle");
        *out = nullptr;
        auto pt = std::make_unique<vkrt_pt>();
        pt->ctx = ctx;
        if(settings) pt->settings = *settings;
        else default_settings(pt->settings);
        check_settings(pt->settings);
        pt->random_engine.seed(seed);
        VKRT_HIP(hipEventCreate(&pt->ev0));
        VKRT_HIP(hipEventCreate(&pt->ev1));
        VKRT_HIP(hipEventCreateWithFlags(&pt->ev_fork, hipEventDisableTiming));
        VKRT_HIP(hipEventCreateWithFlags(&pt->ev_join, hipEventDisableTiming));
        VKRT_HIP(hipStreamCreateWithFlags(&pt->train_stream, hipStreamNonBlocking));
        *out = pt.release();
    });
}

namespace vkrt
{
static void release_scene_accel(vkrt_pt *pt);
}

void vkrt_pt_destroy(vkrt_pt *pt)
{
    if(!pt) return;
    (void)hipSetDevice(pt->ctx->device);
    (void)hipStreamSynchronize(pt->ctx->stream);
    if(pt->train_stream) (void)hipStreamSynchronize(pt->train_stream);
    release_scene_accel(pt);
    if(pt->ev0) (void)hipEventDestroy(pt->ev0);
    if(pt->ev1) (void)hipEventDestroy(pt->ev1);
    if(pt->ev_fork) (void)hipEventDestroy(pt->ev_fork);
    if(pt->ev_join) (void)hipEventDestroy(pt->ev_join);
    if(pt->train_stream) (void)hipStreamDestroy(pt->train_stream);
    delete pt;
}

int vkrt_pt_set_settings(vkrt_pt *pt, const vkrt_pt_settings *settings)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    return guarded(pt->ctx, [&] {
        VKRT_REQUIRE(settings, "null settings");
        check_settings(*settings);
        pt->settings = *settings;
    });
}

int vkrt_pt_get_settings(const vkrt_pt *pt, vkrt_pt_settings *out)
{
    if(!pt || !out) return VKRT_E_INVALID_ARG;
    *out = pt->settings;
    return VKRT_OK;
}

static bool finite3(const float *v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

int vkrt_pt_set_launch_mode(vkrt_pt *pt, int mode, uint32_t tail_depth)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    return guarded(pt->ctx, [&] {
        VKRT_REQUIRE(mode == VKRT_PT_LAUNCH_PERSISTENT || mode == VKRT_PT_LAUNCH_FULL_GRID,
                     "launch mode must be VKRT_PT_LAUNCH_PERSISTENT or VKRT_PT_LAUNCH_FULL_GRID");
        VKRT_REQUIRE(tail_depth >= 1, "tail_depth must be at least 1 (VKRT_PT_NO_TAIL: no tail)");
        pt->launch_mode = mode;
        pt->tail_depth = tail_depth;
    });
}

int vkrt_pt_set_lights(vkrt_pt *pt, const vkrt_light *lights, uint32_t n)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    return guarded(pt->ctx, [&] {
        VKRT_REQUIRE(n == 0 || lights, "null light array");
        VKRT_REQUIRE(n <= 65536u, "at most 65536 lights");
        // the gather of update_trace_descriptors: positive intensity, Area reserved (PBRPathTracer.cpp:630-637)
        std::vector<vkrt_light> kept;
        kept.reserve(n);
        for(uint32_t i = 0; i < n; ++i)
        {
            const vkrt_light &l = lights[i];
            VKRT_REQUIRE(l.type <= VKRT_LIGHT_DISK, "unknown light type");
            if(!(l.intensity > 0.f) || l.type == VKRT_LIGHT_AREA) continue;
            VKRT_REQUIRE(std::isfinite(l.intensity) && finite3(l.color) && finite3(l.position) && finite3(l.direction),
                         "light with a non-finite field");
            const bool body = l.type == VKRT_LIGHT_RECT || l.type == VKRT_LIGHT_SPHERE || l.type == VKRT_LIGHT_TUBE ||
                              l.type == VKRT_LIGHT_DISK;
            VKRT_REQUIRE(!body || (l.size_x > 0.f && std::isfinite(l.size_x)), "light body needs a positive size_x");
            VKRT_REQUIRE(!(l.type == VKRT_LIGHT_RECT || l.type == VKRT_LIGHT_TUBE) ||
                                 (l.size_y > 0.f && std::isfinite(l.size_y)),
                         "rect and tube lights need a positive size_y");
            VKRT_REQUIRE(l.type != VKRT_LIGHT_RECT || finite3(l.tangent), "rect light needs a finite tangent");
            VKRT_REQUIRE(l.type > VKRT_LIGHT_SPOT || l.range > 0.f,
                         "omni and spot lights need range > 0 (infinity: no range window)");
            kept.push_back(l);
        }
        pt->n_lights = 0;
        if(!kept.empty())
        {
            pt->lights.reserve(kept.size() * sizeof(vkrt_light));
            VKRT_HIP(hipMemcpyAsync(pt->lights.ptr, kept.data(), kept.size() * sizeof(vkrt_light),
                                    hipMemcpyHostToDevice, pt->ctx->stream));
            VKRT_HIP(hipStreamSynchronize(pt->ctx->stream));
        }
        pt->n_lights = static_cast<uint32_t>(kept.size());
    });
}

namespace vkrt
{

//! the scene's arrays, materials and textures on the device (the part of set_scene both scene kinds share)
static void upload_scene(vkrt_pt *pt, const vkrt_pt_scene *sc)
{
        VKRT_REQUIRE(sc, "null scene");
        VKRT_REQUIRE(sc->n_tri == 0 || (sc->positions && sc->normals && sc->indices && sc->tri_object),
                     "null scene arrays");
        VKRT_REQUIRE(sc->n_objects > 0 && sc->object_material && sc->n_materials > 0 && sc->materials,
                     "a scene needs objects and materials");
        hipStream_t s = pt->ctx->stream;
        for(uint32_t o = 0; o < sc->n_objects; ++o)
            VKRT_REQUIRE(sc->object_material[o] < sc->n_materials, "object material index out of range");
        std::vector<float> lo(3 * (size_t)sc->n_objects, 3.40282347e+38f), hi(3 * (size_t)sc->n_objects, -3.40282347e+38f);
        for(uint32_t t = 0; t < sc->n_tri; ++t)
        {
            const uint32_t o = sc->tri_object[t];
            VKRT_REQUIRE(o < sc->n_objects, "triangle object index out of range");
            for(int k = 0; k < 3; ++k)
            {
                const uint32_t vi = sc->indices[3 * (size_t)t + k];
                VKRT_REQUIRE(vi < sc->n_vert, "vertex index out of range");
                for(int a = 0; a < 3; ++a)
                {
                    const float c = sc->positions[3 * (size_t)vi + a];
                    lo[3 * o + a] = c < lo[3 * o + a] ? c : lo[3 * o + a];
                    hi[3 * o + a] = hi[3 * o + a] < c ? c : hi[3 * o + a];
                }
            }
        }
        // tmax_inside_volume = 2 * length(scale * aabb.half_extents()) (raypipeline.slang:370), unit scale
        // half extents kept per object: an instance scales them before the length
        std::vector<float> tmax(sc->n_objects), half(3 * (size_t)sc->n_objects);
        for(uint32_t o = 0; o < sc->n_objects; ++o)
        {
            float he[3];
            for(int a = 0; a < 3; ++a) he[a] = hi[3 * o] < lo[3 * o] ? 0.f : (hi[3 * o + a] - lo[3 * o + a]) / 2.f;
            for(int a = 0; a < 3; ++a) half[3 * (size_t)o + a] = he[a];
            tmax[o] = 2.f * std::sqrt((he[0] * he[0] + he[1] * he[1]) + he[2] * he[2]);
        }
        std::vector<gpu_material> mats(sc->n_materials);
        for(uint32_t m = 0; m < sc->n_materials; ++m)
        {
            const vkrt_material &x = sc->materials[m];
            mats[m].color = make_float4(x.color[0], x.color[1], x.color[2], x.color[3]);
            mats[m].emission = make_float4(x.emission[0], x.emission[1], x.emission[2], x.emission[3]);
            mats[m].mrti = make_float4(x.metalness, x.roughness, x.transmission, x.ior);
            mats[m].cc = make_float4(x.clearcoat, x.clearcoat_roughness, x.sheen_roughness, x.diffuse_transmission);
            mats[m].sheen = make_float4(x.sheen_color[0], x.sheen_color[1], x.sheen_color[2], x.two_sided ? 1.f : 0.f);
            mats[m].dtc = make_float4(x.diffuse_transmission_color[0], x.diffuse_transmission_color[1],
                                      x.diffuse_transmission_color[2], 0.f);
            VKRT_REQUIRE(x.color_texture <= sc->n_textures, "material texture index out of range");
            VKRT_REQUIRE(x.alpha_mode <= VKRT_ALPHA_BLEND, "unknown alpha mode");
            uint32_t cutoff_bits;
            std::memcpy(&cutoff_bits, &x.alpha_cutoff, 4);
            // MASK and BLEND drive the any-hit test
            mats[m].tex = make_uint4(x.color_texture, x.alpha_mode, cutoff_bits, 0u);
            const uint32_t slots[6] = {x.normal_texture, x.emission_texture, x.ao_rough_metal_texture,
                                       x.transmission_texture, x.diffuse_transmission_texture,
                                       x.diffuse_transmission_color_texture};
            for(uint32_t k: slots) VKRT_REQUIRE(k <= sc->n_textures, "material texture index out of range");
            mats[m].tex2 = make_uint4(slots[0], slots[1], slots[2], slots[3]);
            mats[m].tex3 = make_uint4(slots[4], slots[5], 0u, 0u);
            // KHR_materials_volume: sigma_t = -log(attenuation color) / attenuation distance (raypipeline.slang:493)
            const float ad = x.attenuation_distance;
            VKRT_REQUIRE(std::isfinite(ad) && ad >= 0.f, "attenuation_distance must be finite and >= 0 (0: no volume)");
            mats[m].vol = make_float4(0.f, 0.f, 0.f, 0.f);
            if(ad > 0.f)
            {
                float sg[3];
                for(int k = 0; k < 3; ++k)
                {
                    const float c = x.attenuation_color[k];
                    VKRT_REQUIRE(c > 0.f && c <= 1.f, "attenuation_color must lie in (0, 1]");
                    sg[k] = -std::log(c) / ad;
                }
                mats[m].vol = make_float4(sg[0], sg[1], sg[2], 1.f);
            }
        }
        // texture coordinates (zeros when the scene has none) and the textures' mip chains: each level halves
        // the previous one (at least 1 texel) with a 2 x 2 box filter, coordinates clamped at odd edges,
        // rounded to nearest ((sum + 2) / 4)
        std::vector<float> uv(2 * (size_t)sc->n_vert, 0.f), tan(3 * (size_t)sc->n_vert, 0.f);
        if(sc->texcoords) std::memcpy(uv.data(), sc->texcoords, uv.size() * 4);
        if(sc->tangents) std::memcpy(tan.data(), sc->tangents, tan.size() * 4);
        VKRT_REQUIRE(sc->n_textures == 0 || sc->textures, "null texture array");
        std::vector<tex_desc> descs(sc->n_textures);
        std::vector<uint32_t> texels;
        for(uint32_t t = 0; t < sc->n_textures; ++t)
        {
            const vkrt_texture &tx = sc->textures[t];
            VKRT_REQUIRE(tx.rgba && tx.width > 0 && tx.height > 0 && tx.width <= 16384 && tx.height <= 16384,
                         "textures need pixels and sizes in 1..16384");
            tex_desc &d = descs[t];
            d = tex_desc{};
            d.width = tx.width;
            d.height = tx.height;
            VKRT_REQUIRE(tx.address_mode_u <= VKRT_ADDRESS_CLAMP_TO_EDGE && tx.address_mode_v <= VKRT_ADDRESS_CLAMP_TO_EDGE &&
                                 tx.min_filter <= VKRT_FILTER_NEAREST && tx.mag_filter <= VKRT_FILTER_NEAREST,
                         "unknown texture address mode or filter");
            d.sampler = tx.address_mode_u | (tx.address_mode_v << 2) | (tx.min_filter << 4) | (tx.mag_filter << 5);
            uint32_t w = tx.width, h = tx.height;
            d.offset[0] = static_cast<uint32_t>(texels.size());
            for(size_t i = 0; i < (size_t)w * h; ++i)
            {
                uint32_t c;
                std::memcpy(&c, tx.rgba + 4 * i, 4);
                texels.push_back(c);
            }
            d.levels = 1;
            while((w > 1 || h > 1) && d.levels < MAX_TEX_LEVELS)
            {
                const uint32_t pw = w, ph = h, src = d.offset[d.levels - 1];
                w = std::max(1u, w / 2);
                h = std::max(1u, h / 2);
                d.offset[d.levels] = static_cast<uint32_t>(texels.size());
                for(uint32_t y = 0; y < h; ++y)
                    for(uint32_t x = 0; x < w; ++x)
                    {
                        const uint32_t xa = std::min(2 * x, pw - 1), xb = std::min(2 * x + 1, pw - 1);
                        const uint32_t ya = std::min(2 * y, ph - 1), yb = std::min(2 * y + 1, ph - 1);
                        const uint32_t q[4] = {texels[src + ya * pw + xa], texels[src + ya * pw + xb],
                                               texels[src + yb * pw + xa], texels[src + yb * pw + xb]};
                        uint32_t out = 0;
                        for(int ch = 0; ch < 4; ++ch)
                        {
                            uint32_t sum = 2;
                            for(uint32_t qq: q) sum += (qq >> (8 * ch)) & 0xFFu;
                            out |= (sum / 4) << (8 * ch);
                        }
                        texels.push_back(out);
                    }
                ++d.levels;
            }
            VKRT_REQUIRE(texels.size() < (1ull << 31), "textures too large");
        }
        auto upload = [&](device_buffer &buf, const void *src, size_t bytes) {
            buf.reserve(bytes);
            if(bytes) VKRT_HIP(hipMemcpyAsync(buf.ptr, src, bytes, hipMemcpyHostToDevice, s));
        };
        upload(pt->pos, sc->positions, (size_t)sc->n_vert * 12);
        upload(pt->nrm, sc->normals, (size_t)sc->n_vert * 12);
        upload(pt->idx, sc->indices, (size_t)sc->n_tri * 12);
        upload(pt->tri_object, sc->tri_object, (size_t)sc->n_tri * 4);
        upload(pt->object_material, sc->object_material, (size_t)sc->n_objects * 4);
        upload(pt->object_tmax, tmax.data(), tmax.size() * 4);
        upload(pt->object_he, half.data(), half.size() * 4);
        upload(pt->mats, mats.data(), mats.size() * sizeof(gpu_material));
        pt->any_volume = std::any_of(mats.begin(), mats.end(), [](const gpu_material &m) { return m.vol.w > 0.f; });
        VKRT_REQUIRE(!pt->any_volume || sc->n_materials < NO_MEDIUM,
                     "scenes with interior media take at most 65,534 materials");
        pt->any_alpha = std::any_of(mats.begin(), mats.end(),
                                    [](const gpu_material &m) { return m.tex.y != VKRT_ALPHA_OPAQUE; });
        upload(pt->uv, uv.data(), uv.size() * 4);
        upload(pt->tan, tan.data(), tan.size() * 4);
        upload(pt->tex_desc_buf, descs.data(), descs.size() * sizeof(tex_desc));
        upload(pt->texels, texels.data(), texels.size() * 4);
        VKRT_HIP(hipStreamSynchronize(s));
        pt->n_tri = sc->n_tri;
        pt->n_vert = sc->n_vert;
        pt->n_objects = sc->n_objects;
        pt->n_materials = sc->n_materials;
        pt->environment = sc->environment;
}

//! the 3 x 4 object-to-world matrix of an instance: columns R(q) (S e_j), the last column T, the map
//! apply_transform computes (transform.hpp: rotate(q, p * S) + T), evaluated in double and rounded once
static void instance_matrix(const vkrt_pt_instance &in, float m[12])
{
    const double qx = in.rotation[0], qy = in.rotation[1], qz = in.rotation[2], qw = in.rotation[3];
    for(int j = 0; j < 3; ++j)
    {
        double v[3] = {0.0, 0.0, 0.0};
        v[j] = in.scale[j];
        // v + 2 q.xyz x (q.xyz x v + w v)
        const double c1[3] = {qy * v[2] - qz * v[1] + qw * v[0], qz * v[0] - qx * v[2] + qw * v[1],
                              qx * v[1] - qy * v[0] + qw * v[2]};
        const double c2[3] = {qy * c1[2] - qz * c1[1], qz * c1[0] - qx * c1[2], qx * c1[1] - qy * c1[0]};
        for(int a = 0; a < 3; ++a) m[4 * a + j] = static_cast<float>(v[a] + 2.0 * c2[a]);
    }
    for(int a = 0; a < 3; ++a) m[4 * a + 3] = in.translation[a];
}

//! frees the acceleration structures of the current scene (either kind)
static void release_scene_accel(vkrt_pt *pt)
{
    if(pt->scene) vkrt_scene_destroy(pt->scene);
    pt->scene = nullptr;
    for(vkrt_bvh *b: pt->mesh_bvhs) vkrt_bvh_destroy(b);
    pt->mesh_bvhs.clear();
    if(pt->bvh) vkrt_bvh_destroy(pt->bvh);
    pt->bvh = nullptr;
    pt->instanced = false;
    pt->tlas = pt_tlas{};
}

//! a PLOC-built BVH (lower SAH cost than the LBVH, 10-35 % faster traversal) over triangles [first, first + n)
//! of the uploaded scene, its leaf records tagged with materials, flags and the scene's triangle numbers
static vkrt_bvh *build_tagged_bvh(vkrt_pt *pt, uint32_t first, uint32_t n)
{
    vkrt_bvh_params params{};
    params.builder = VKRT_BUILDER_PLOC;
    params.max_leaf_size = 2;
    vkrt_bvh *bvh = nullptr;
    const int rc = vkrt_bvh_build(pt->ctx, pt->pos.as<float>(), pt->n_vert, 12, pt->idx.as<uint32_t>() + 3 * (size_t)first,
                                  n, &params, VKRT_DEVICE_PTRS, &bvh);
    if(rc != VKRT_OK) throw error(rc, "BVH build failed: " + pt->ctx->last_error);
    if(n)
        k_pt_tag<<<div_up(n, 256), 256, 0, pt->ctx->stream>>>(bvh->tris.as<float4>(), n,
                                                               pt->tri_object.as<uint32_t>(),
                                                               pt->object_material.as<uint32_t>(),
                                                               pt->mats.as<gpu_material>(), first);
    VKRT_HIP(hipGetLastError());
    return bvh;
}

}// namespace vkrt

int vkrt_pt_set_scene(vkrt_pt *pt, const vkrt_pt_scene *sc)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    return guarded(pt->ctx, [&] {
        upload_scene(pt, sc);
        release_scene_accel(pt);
        pt->bvh = build_tagged_bvh(pt, 0, sc->n_tri);
        VKRT_HIP(hipStreamSynchronize(pt->ctx->stream));
        for(int k = 0; k < 6; ++k) pt->bounds[k] = pt->bvh->info.bounds[k];
        fit_cache_bounds(pt);
    });
}

int vkrt_pt_set_instanced_scene(vkrt_pt *pt, const vkrt_pt_scene *sc, const uint32_t *mesh_first, uint32_t n_meshes,
                                const vkrt_pt_instance *instances, uint32_t n_instances)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    return guarded(pt->ctx, [&] {
        VKRT_REQUIRE(sc, "null scene");
        VKRT_REQUIRE(n_meshes > 0 && mesh_first, "an instanced scene needs meshes");
        VKRT_REQUIRE(n_instances == 0 || instances, "null instances");
        VKRT_REQUIRE(mesh_first[0] == 0 && mesh_first[n_meshes] == sc->n_tri, "mesh ranges must cover the triangles");
        for(uint32_t m = 0; m < n_meshes; ++m)
            VKRT_REQUIRE(mesh_first[m] <= mesh_first[m + 1], "mesh ranges must not decrease");
        for(uint32_t i = 0; i < n_instances; ++i)
        {
            const vkrt_pt_instance &in = instances[i];
            VKRT_REQUIRE(in.mesh < n_meshes, "instance mesh index out of range");
            bool finite = true;
            for(int k = 0; k < 3; ++k)
                finite = finite && std::isfinite(in.translation[k]) && std::isfinite(in.scale[k]) && in.scale[k] != 0.f;
            for(int k = 0; k < 4; ++k) finite = finite && std::isfinite(in.rotation[k]);
            VKRT_REQUIRE(finite, "instance transforms must be finite with non-zero scale");
        }
        upload_scene(pt, sc);
        release_scene_accel(pt);
        hipStream_t s = pt->ctx->stream;
        for(uint32_t m = 0; m < n_meshes; ++m)
            pt->mesh_bvhs.push_back(build_tagged_bvh(pt, mesh_first[m], mesh_first[m + 1] - mesh_first[m]));
        VKRT_HIP(hipStreamSynchronize(s));
        // the instances' matrices (object to world = T R S, the map apply_transform computes), their shading
        // transforms, and the world box
        std::vector<vkrt_instance> insts(n_instances);
        std::vector<float4> xf(3 * (size_t)n_instances);
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for(uint32_t i = 0; i < n_instances; ++i)
        {
            const vkrt_pt_instance &in = instances[i];
            float m12[12];
            instance_matrix(in, m12);
            insts[i] = vkrt_instance{};
            vkrt_instance_set_transform(&insts[i], m12);
            insts[i].blas = in.mesh;
            insts[i].custom_index = i;
            xf[3 * i] = make_float4(in.rotation[0], in.rotation[1], in.rotation[2], in.rotation[3]);
            xf[3 * i + 1] = make_float4(in.translation[0], in.translation[1], in.translation[2], in.scale[0]);
            xf[3 * i + 2] = make_float4(in.scale[1], in.scale[2], 0.f, 0.f);
            const vkrt_bvh *b = pt->mesh_bvhs[in.mesh];
            if(b->empty) continue;
            const float *bb = b->info.bounds;
            for(int c = 0; c < 8; ++c)
            {
                const double q[3] = {(c & 1) ? bb[3] : bb[0], (c & 2) ? bb[4] : bb[1], (c & 4) ? bb[5] : bb[2]};
                for(int a = 0; a < 3; ++a)
                {
                    const double w = (double)m12[4 * a] * q[0] + (double)m12[4 * a + 1] * q[1] +
                                     (double)m12[4 * a + 2] * q[2] + (double)m12[4 * a + 3];
                    lo[a] = std::min(lo[a], w);
                    hi[a] = std::max(hi[a], w);
                }
            }
        }
        const int rc = vkrt_scene_create(pt->ctx, pt->mesh_bvhs.data(), n_meshes, insts.data(), n_instances,
                                         &pt->scene);
        if(rc != VKRT_OK) throw error(rc, "instanced scene: " + pt->ctx->last_error);
        const scene_view v = scene_device_view(pt->scene);
        pt->tlas = pt_tlas{v.tlas, v.root, v.empty, v.order, v.insts, v.blases};
        pt->inst_xf.reserve(std::max<size_t>(1, xf.size()) * 16);
        if(!xf.empty())
            VKRT_HIP(hipMemcpyAsync(pt->inst_xf.ptr, xf.data(), xf.size() * 16, hipMemcpyHostToDevice, s));
        VKRT_HIP(hipStreamSynchronize(s));
        for(int a = 0; a < 3; ++a)
        {
            pt->bounds[a] = static_cast<float>(lo[a]);
            pt->bounds[a + 3] = static_cast<float>(hi[a]);
        }
        pt->instanced = true;
        fit_cache_bounds(pt);
    });
}

int vkrt_pt_trace_transmittance(vkrt_pt *pt, const vkrt_ray *rays, uint32_t n, float *out4, int flags)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    vkrt_ctx *ctx = pt->ctx;
    return guarded(ctx, [&] {
        VKRT_REQUIRE(pt->bvh || pt->instanced, "vkrt_pt_set_scene first");
        VKRT_REQUIRE(n == 0 || (rays && out4), "null array argument");
        if(!n) return;
        const vkrt_ray *dr = input(ctx, 1, rays, n, flags);
        float4 *dout = reinterpret_cast<float4 *>(output(ctx, 2, out4, (size_t)n * 4, flags));
        const pt_scene_dev sc = scene_dev(pt);
        const alpha_filter flt{sc.idx, sc.uv, sc.mats, sc.tex, sc.texels, false};
        const bool cull = (flags & VKRT_CULL_BACK) != 0;
        const float4 *r4 = reinterpret_cast<const float4 *>(dr);
        if(pt->instanced)
            k_pt_transmittance<true><<<div_up(n, PT_BLOCK), PT_BLOCK, 0, ctx->stream>>>(
                    nullptr, 0, nullptr, 0, false, true, nullptr, pt->mats.as<gpu_material>(), r4, n, cull, dout, flt,
                    pt->tlas);
        else
        {
            const vkrt_bvh *b = pt->bvh;
            k_pt_transmittance<false><<<div_up(n, PT_BLOCK), PT_BLOCK, 0, ctx->stream>>>(
                    b->nodes.as<vkrt_bvh_node>(), b->root_ref, b->nodes4.as<vkrt_bvh4_node>(), b->root4, b->wide,
                    b->empty, b->tris.as<float4>(), pt->mats.as<gpu_material>(), r4, n, cull, dout, flt, pt->tlas);
        }
        VKRT_HIP(hipGetLastError());
        finish_output(ctx, 2, out4, (size_t)n * 4, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_denoise(vkrt_ctx *ctx, const vkrt_pixel *pixels, uint32_t width, uint32_t height, uint32_t denoise,
                 float *out_rgba, int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(width > 0 && height > 0 && (uint64_t)width * height <= (1u << 26), "bad image size");
        VKRT_REQUIRE(pixels && out_rgba, "null array argument");
        const size_t n = (size_t)width * height;
        const vkrt_pixel *dp = input(ctx, 0, pixels, n, flags);
        ctx->scratch[2].reserve(2 * n * sizeof(float4));
        float4 *ping = ctx->scratch[2].as<float4>(), *pong = ping + n;
        float4 *res = run_denoise(ctx->stream, dp, width, height, denoise != 0, make_tonemap(false, 0.f, 1.f), ping,
                                  pong);
        VKRT_HIP(hipMemcpyAsync(out_rgba, res, n * sizeof(float4),
                                (flags & VKRT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                ctx->stream));
        finish_call(ctx, flags);
    });
}

int vkrt_tonemap(vkrt_ctx *ctx, const float *in_rgba, uint32_t n, float exposure, float gamma, float *out_rgba,
                 int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(n <= (1u << 26), "too many texels");
        VKRT_REQUIRE(n == 0 || (in_rgba && out_rgba), "null array argument");
        const tonemap_params tm = make_tonemap(true, exposure, gamma);
        if(n == 0) return;
        const float4 *din = input(ctx, 0, reinterpret_cast<const float4 *>(in_rgba), n, flags);
        float4 *dout = output<float4>(ctx, 1, reinterpret_cast<float4 *>(out_rgba), n, flags);
        k_tonemap<<<div_up(n, 256u), 256, 0, ctx->stream>>>(din, dout, n, tm);
        VKRT_HIP(hipGetLastError());
        finish_output(ctx, 1, reinterpret_cast<float4 *>(out_rgba), n, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_pt_read_image(vkrt_pt *pt, float *out_rgba, int flags)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    vkrt_ctx *ctx = pt->ctx;
    return guarded(ctx, [&] {
        VKRT_REQUIRE(out_rgba, "null output");
        VKRT_REQUIRE(pt->width > 0, "nothing rendered yet");
        const size_t n = (size_t)pt->width * pt->height;
        pt->image.reserve(2 * n * sizeof(float4));
        float4 *ping = pt->image.as<float4>(), *pong = ping + n;
        const tonemap_params tm = make_tonemap(pt->settings.tonemap != 0, pt->settings.exposure, pt->settings.gamma);
        float4 *res = run_denoise(ctx->stream, pt->pixels.as<vkrt_pixel>(), pt->width, pt->height,
                                  pt->settings.denoising != 0, tm, ping, pong);
        VKRT_HIP(hipMemcpyAsync(out_rgba, res, n * sizeof(float4),
                                (flags & VKRT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                ctx->stream));
        finish_call(ctx, flags);
    });
}

void vkrt_pt_default_cache_config(vkrt_pt_cache_config *out)
{
    if(!out) return;
    *out = {};
    out->cache_depth = 2;
    out->train_stride = 64;
    out->max_train_records = 16384;
}

int vkrt_pt_set_radiance_cache(vkrt_pt *pt, vkrt_rc *rc, const vkrt_pt_cache_config *cfg)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    return guarded(pt->ctx, [&] {
        vkrt_pt_cache_config c{};
        if(cfg) c = *cfg;
        else vkrt_pt_default_cache_config(&c);
        VKRT_REQUIRE(c.cache_depth >= 1 && c.cache_depth <= 64, "cache_depth must be in 1..64");
        VKRT_REQUIRE(c.train_stride >= 1 && c.train_stride <= (1u << 20), "train_stride must be in 1..2^20");
        VKRT_REQUIRE(c.max_train_records <= (1u << 22), "max_train_records must be at most 2^22");
        VKRT_REQUIRE(!rc || rc->ctx == pt->ctx, "the cache belongs to another context");
        pt->rc = rc;
        pt->cache_cfg = c;
        fit_cache_bounds(pt);
    });
}

int vkrt_pt_render(vkrt_pt *pt, const vkrt_camera *camera)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    return guarded(pt->ctx, [&] {
        VKRT_REQUIRE(camera, "null camera");
        VKRT_REQUIRE(pt->bvh || pt->instanced, "vkrt_pt_set_scene first");
        const vkrt_pt_settings settings = pt->settings;// settings are copied per frame (PBRPathTracer.cpp:198)
        resize_storage(pt, settings.resolution[0], settings.resolution[1]);
        pt->stats.traced = 0;
        if(!settings.max_num_batches || !settings.suspend_trace_when_done || pt->batch_index < settings.max_num_batches)
        {
            const uint32_t seed = static_cast<uint32_t>(pt->random_engine());
            trace_batch(pt, *camera, seed);
            pt->batch_index = std::min<uint32_t>(pt->batch_index + 1, settings.max_num_batches);
        }
    });
}

int vkrt_pt_reset_accumulator(vkrt_pt *pt)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    if(!pt->settings.suppress_reset) pt->batch_index = 0;
    return VKRT_OK;
}

uint32_t vkrt_pt_current_batch(const vkrt_pt *pt) { return pt ? pt->batch_index : 0u; }

int vkrt_pt_read_pixels(vkrt_pt *pt, vkrt_pixel *out, int flags)
{
    if(!pt) return VKRT_E_INVALID_ARG;
    return guarded(pt->ctx, [&] {
        VKRT_REQUIRE(out, "null output");
        VKRT_REQUIRE(pt->width > 0, "nothing rendered yet");
        const size_t bytes = (size_t)pt->width * pt->height * sizeof(vkrt_pixel);
        VKRT_HIP(hipMemcpyAsync(out, pt->pixels.ptr, bytes,
                                (flags & VKRT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                pt->ctx->stream));
        VKRT_HIP(hipStreamSynchronize(pt->ctx->stream));
    });
}

int vkrt_pt_get_stats(const vkrt_pt *pt, vkrt_pt_stats *out)
{
    if(!pt || !out) return VKRT_E_INVALID_ARG;
    *out = pt->stats;
    return VKRT_OK;
}

}// extern "C"
