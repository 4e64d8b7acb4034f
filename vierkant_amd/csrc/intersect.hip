// intersect.hip - batched versions of vierkant's intersect(...) overloads and the brute-force
// closest/any-hit loop, for gfx950.
//
// Pairwise kernels evaluate the reference arithmetic (vierkant/rt_math.hpp) one element per lane.
// The brute-force kernels tile (rays x triangle chunks): each wave64 holds 64 rays, walks one
// triangle chunk whose addresses are wave-uniform (scalar loads feed the VALU directly), and merges
// its best hit into a per-ray 64-bit key (ordered t << 32 | prim) with one atomic min, so ties on t
// resolve to the lowest triangle index no matter which chunk finishes first.
#include <cfloat>
#include <mutex>

#include "context.hpp"
#include "kernels.hpp"
#include "vierkant/rt_math.hpp"

using namespace vierkant::rt;

namespace vkrt
{

__device__ __forceinline__ vec3 ld3(const float *p) { return vec3{p[0], p[1], p[2]}; }

__device__ __forceinline__ void load_ray(const float4 *__restrict__ rays, uint32_t i, vec3 &o, float &tmin, vec3 &d,
                                         float &tmax)
{
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    o = vec3{a.x, a.y, a.z};
    tmin = a.w;
    d = vec3{b.x, b.y, b.z};
    tmax = b.w;
}

// ---------------------------------------------------------------------------------------------
// pairwise tests
// ---------------------------------------------------------------------------------------------
__global__ void k_pair_tri(const float *__restrict__ tris, const float4 *__restrict__ rays, uint32_t n,
                           vkrt_tri_result *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    vec3 o, d;
    float tmin, tmax;
    load_ray(rays, i, o, tmin, d, tmax);
    const float *t9 = tris + 9 * (size_t)i;
    vec3 v0 = ld3(t9), v1 = ld3(t9 + 3), v2 = ld3(t9 + 6);
    float t = 0.f, u = 0.f, v = 0.f;
    bool hit = mt_reference(o, d, v0, v1 - v0, v2 - v0, t, u, v);
    out[i] = hit ? vkrt_tri_result{INTERSECT, t, u, v} : vkrt_tri_result{REJECT, 0.f, 0.f, 0.f};
}

__global__ void k_pair_aabb(const float *__restrict__ boxes, const float4 *__restrict__ rays, uint32_t n,
                            vkrt_result *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    vec3 o, d;
    float tmin, tmax;
    load_ray(rays, i, o, tmin, d, tmax);
    const float *b = boxes + 6 * (size_t)i;
    float dist = 0.f;
    uint32_t type = aabb_reference(ld3(b), ld3(b + 3), o, d, dist);
    out[i] = vkrt_result{type, type ? dist : 0.f};
}

__global__ void k_pair_obb(const float *__restrict__ obbs, const float4 *__restrict__ rays, uint32_t n,
                           vkrt_result *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    vec3 o, d;
    float tmin, tmax;
    load_ray(rays, i, o, tmin, d, tmax);
    const float *b = obbs + 15 * (size_t)i;
    const vec3 axis[3] = {ld3(b + 3), ld3(b + 6), ld3(b + 9)};
    float dist = 0.f;
    uint32_t type = obb_reference(ld3(b), axis, ld3(b + 12), o, d, dist);
    out[i] = vkrt_result{type, type ? dist : 0.f};
}

__global__ void k_pair_sphere(const float *__restrict__ spheres, const float4 *__restrict__ rays, uint32_t n,
                              vkrt_result *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    vec3 o, d;
    float tmin, tmax;
    load_ray(rays, i, o, tmin, d, tmax);
    const float *s = spheres + 4 * (size_t)i;
    float dist = 0.f;
    uint32_t type = sphere_reference(ld3(s), s[3], o, d, dist);
    out[i] = vkrt_result{type, type ? dist : 0.f};
}

__global__ void k_pair_plane(const float *__restrict__ planes, const float4 *__restrict__ rays, uint32_t n,
                             vkrt_result *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    vec3 o, d;
    float tmin, tmax;
    load_ray(rays, i, o, tmin, d, tmax);
    const float *p = planes + 4 * (size_t)i;
    const float c[4] = {p[0], p[1], p[2], p[3]};
    float dist = 0.f;
    uint32_t type = plane_reference(c, o, d, dist);
    out[i] = vkrt_result{type, type ? dist : 0.f};
}

__global__ void k_pair_tri_box(const float *__restrict__ tris, const float *__restrict__ boxes, uint32_t n,
                               uint32_t *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    float t[9], b[6];
    for(int k = 0; k < 9; ++k) t[k] = tris[9 * (size_t)i + k];
    for(int k = 0; k < 6; ++k) b[k] = boxes[6 * (size_t)i + k];
    out[i] = tri_box_overlap(t, b, b + 3);
}

__global__ void k_pair_tri_tri(const float *__restrict__ ta, const float *__restrict__ tb, uint32_t n,
                               uint32_t *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    float a[9], b[9];
    for(int k = 0; k < 9; ++k) a[k] = ta[9 * (size_t)i + k], b[k] = tb[9 * (size_t)i + k];
    out[i] = tri_tri_overlap(a, b);
}

// ---------------------------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------------------------
__global__ void k_fill_u64(uint64_t *__restrict__ p, uint32_t n, uint64_t value)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) p[i] = value;
}

__global__ void k_fill_u8(uint8_t *__restrict__ p, uint32_t n, uint8_t value)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) p[i] = value;
}

//! (v0, v1, v2) -> (v0, e1 = v1 - v0, e2 = v2 - v0): the first line of the reference test
__global__ void k_tri_prep(const float *__restrict__ tris, uint32_t n, float *__restrict__ pre)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const float *t = tris + 9 * (size_t)i;
    float *p = pre + 9 * (size_t)i;
    vec3 v0 = ld3(t), e1 = ld3(t + 3) - v0, e2 = ld3(t + 6) - v0;
    p[0] = v0.x, p[1] = v0.y, p[2] = v0.z;
    p[3] = e1.x, p[4] = e1.y, p[5] = e1.z;
    p[6] = e2.x, p[7] = e2.y, p[8] = e2.z;
}

// ---------------------------------------------------------------------------------------------
// brute force
// ---------------------------------------------------------------------------------------------
// 65,536 rays x 20,480 triangles: 1.18 ms per launch against 1.23 one triangle at a time; eight at a time
// (98 VGPRs) is no faster
constexpr int BRUTE_UNROLL = 4;

__global__ void __launch_bounds__(256) k_brute_closest(const float4 *__restrict__ rays, uint32_t n_ray,
                                                       const float *__restrict__ pre, uint32_t n_tri,
                                                       uint32_t chunk, uint64_t *__restrict__ keys)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t t0 = blockIdx.y * chunk;
    const uint32_t t1 = min(n_tri, t0 + chunk);
    vec3 o{0.f, 0.f, 0.f}, d{0.f, 0.f, 1.f};
    float tmin = 0.f, tmax = -1.f;// an empty interval for lanes past the end
    if(r < n_ray) load_ray(rays, r, o, tmin, d, tmax);
    float best = tmax;
    uint32_t best_i = VKRT_INVALID_PRIM;
    // the triangle index is wave-uniform, so records come through scalar loads, four at a time: one wider
    // request and one loop step per four tests (the loop's scalar work competes with the tests for issue)
    uint32_t i = t0;
    for(; i + BRUTE_UNROLL <= t1; i += BRUTE_UNROLL)
    {
        const float *p = pre + 9 * (size_t)i;
        float rec[9 * BRUTE_UNROLL];
#pragma unroll
        for(int k = 0; k < 9 * BRUTE_UNROLL; ++k) rec[k] = p[k];
#pragma unroll
        for(int k = 0; k < BRUTE_UNROLL; ++k)
        {
            const float *q = rec + 9 * k;
            float t, u, v;
            if(mt_fast_window(o, d, vec3{q[0], q[1], q[2]}, vec3{q[3], q[4], q[5]}, vec3{q[6], q[7], q[8]}, tmin,
                              best, t, u, v))
            {
                best = t;
                best_i = i + k;
            }
        }
    }
    for(; i < t1; ++i)
    {
        const float *p = pre + 9 * (size_t)i;
        float t, u, v;
        if(mt_fast_window(o, d, ld3(p), ld3(p + 3), ld3(p + 6), tmin, best, t, u, v))
        {
            best = t;
            best_i = i;
        }
    }
    if(best_i != VKRT_INVALID_PRIM) atomicMin((unsigned long long *)&keys[r], hit_key(best, best_i));
}

__global__ void k_brute_finalize(const float4 *__restrict__ rays, uint32_t n_ray, const float *__restrict__ pre,
                                 const uint64_t *__restrict__ keys, vkrt_hit *__restrict__ hits)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if(r >= n_ray) return;
    uint64_t key = keys[r];
    vkrt_hit h{__int_as_float(0x7f800000), 0.f, 0.f, VKRT_INVALID_PRIM};
    if(key != ~0ull)
    {
        uint32_t prim = static_cast<uint32_t>(key & 0xffffffffu);
        vec3 o, d;
        float tmin, tmax;
        load_ray(rays, r, o, tmin, d, tmax);
        const float *p = pre + 9 * (size_t)prim;
        float t = 0.f, u = 0.f, v = 0.f;
        mt_reference(o, d, ld3(p), ld3(p + 3), ld3(p + 6), t, u, v);
        h = vkrt_hit{t, u, v, prim};
    }
    hits[r] = h;
}

__global__ void __launch_bounds__(256) k_brute_any(const float4 *__restrict__ rays, uint32_t n_ray,
                                                   const float *__restrict__ pre, uint32_t n_tri, uint32_t chunk,
                                                   uint8_t *__restrict__ occluded)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t t0 = blockIdx.y * chunk;
    const uint32_t t1 = min(n_tri, t0 + chunk);
    vec3 o{0.f, 0.f, 0.f}, d{0.f, 0.f, 1.f};
    float tmin = 0.f, tmax = -1.f;
    if(r < n_ray) load_ray(rays, r, o, tmin, d, tmax);
    int occ = (r < n_ray) ? 0 : 1;
    // four records per loop step, as in k_brute_closest; the wave leaves once every lane is occluded
    uint32_t i = t0;
    for(; i + BRUTE_UNROLL <= t1; i += BRUTE_UNROLL)
    {
        if(__all(occ)) break;
        const float *p = pre + 9 * (size_t)i;
        float rec[9 * BRUTE_UNROLL];
#pragma unroll
        for(int k = 0; k < 9 * BRUTE_UNROLL; ++k) rec[k] = p[k];
#pragma unroll
        for(int k = 0; k < BRUTE_UNROLL; ++k)
        {
            const float *q = rec + 9 * k;
            float t, u, v;
            if(!occ && mt_fast_window(o, d, vec3{q[0], q[1], q[2]}, vec3{q[3], q[4], q[5]}, vec3{q[6], q[7], q[8]},
                                      tmin, tmax, t, u, v))
                occ = 1;
        }
    }
    for(; i < t1; ++i)
    {
        if(__all(occ)) break;
        const float *p = pre + 9 * (size_t)i;
        float t, u, v;
        if(!occ && mt_fast_window(o, d, ld3(p), ld3(p + 3), ld3(p + 6), tmin, tmax, t, u, v)) occ = 1;
    }
    if(r < n_ray && occ) occluded[r] = 1;
}

// ---------------------------------------------------------------------------------------------
// compute_aabb: per axis the first point attaining the min (max), as the sequential glm::min/max
// fold of src/intersection.cpp:100-104 keeps the earlier of equal values; NaN never wins a compare.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x)
{
    for(int off = 32; off > 0; off >>= 1)
    {
        uint64_t y = __shfl_xor(x, off);
        x = y < x ? y : x;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x)
{
    for(int off = 32; off > 0; off >>= 1)
    {
        uint64_t y = __shfl_xor(x, off);
        x = y > x ? y : x;
    }
    return x;
}

__global__ void __launch_bounds__(256) k_aabb_reduce(const float *__restrict__ pts, uint32_t n,
                                                     uint64_t *__restrict__ keys)
{
    uint64_t mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0, 0, 0};
    for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        for(int a = 0; a < 3; ++a)
        {
            float c = pts[3 * (size_t)i + a];
            if(c != c) continue;
            uint64_t ord = ordered_bits(c + 0.0f);
            uint64_t kmin = (ord << 32) | i, kmax = (ord << 32) | (0xffffffffu - i);
            mn[a] = kmin < mn[a] ? kmin : mn[a];
            mx[a] = kmax > mx[a] ? kmax : mx[a];
        }
    }
    for(int a = 0; a < 3; ++a)
    {
        uint64_t lo = wave_min_u64(mn[a]), hi = wave_max_u64(mx[a]);
        if((threadIdx.x & 63) == 0)
        {
            if(lo != ~0ull) atomicMin((unsigned long long *)&keys[a], lo);
            if(hi != 0) atomicMax((unsigned long long *)&keys[3 + a], hi);
        }
    }
}

__global__ void k_aabb_finalize(const float *__restrict__ pts, const uint64_t *__restrict__ keys,
                                float *__restrict__ out6)
{
    int a = threadIdx.x;
    if(a >= 6) return;
    const float init = a < 3 ? FLT_MAX : -FLT_MAX;
    uint64_t k = keys[a];
    float r = init;
    if(a < 3 && k != ~0ull)
    {
        float v = pts[3 * (size_t)(k & 0xffffffffu) + a];
        r = v < init ? v : init;
    }
    else if(a >= 3 && k != 0)
    {
        float v = pts[3 * (size_t)(0xffffffffu - (k & 0xffffffffu)) + (a - 3)];
        r = init < v ? v : init;
    }
    out6[a] = r;
}

//! compute_bounding_sphere (src/intersection.cpp:25-92): the extremum indices come from k_aabb_reduce's
//! keys (first occurrence of the strict minimum / maximum, NaN never wins), then one workgroup seeds the
//! sphere from the axis with the largest squared extent and grows it in index order. Per chunk of 1024
//! points the block finds the first point outside the current sphere, that lane applies the reference's
//! update, and the scan resumes after it: the exact sequential result, in parallel within each chunk.
constexpr int SPHERE_BLOCK = 1024;

__global__ void __launch_bounds__(SPHERE_BLOCK) k_sphere_grow(const float *__restrict__ pts, uint32_t n,
                                                             const uint64_t *__restrict__ keys,
                                                             float *__restrict__ out4)
{
    __shared__ float sph[4];
    __shared__ uint32_t first;
    const uint32_t tid = threadIdx.x;
    if(tid == 0)
    {
        uint32_t pmin[3], pmax[3];
        for(int a = 0; a < 3; ++a)
        {
            const float v0 = pts[a];
            const bool nan0 = v0 != v0;// a NaN first coordinate is never replaced
            pmin[a] = (nan0 || keys[a] == ~0ull) ? 0u : static_cast<uint32_t>(keys[a] & 0xffffffffu);
            pmax[a] = (nan0 || keys[3 + a] == 0ull) ? 0u : 0xffffffffu - static_cast<uint32_t>(keys[3 + a] & 0xffffffffu);
        }
        float best = 0.f;
        int axis = 0;
        for(int a = 0; a < 3; ++a)
        {
            const float *p1 = pts + 3 * (size_t)pmin[a], *p2 = pts + 3 * (size_t)pmax[a];
            const float dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
            const float d2 = (dx * dx + dy * dy) + dz * dz;
            if(d2 > best)
            {
                best = d2;
                axis = a;
            }
        }
        const float *p1 = pts + 3 * (size_t)pmin[axis], *p2 = pts + 3 * (size_t)pmax[axis];
        sph[0] = (p1[0] + p2[0]) / 2.f;
        sph[1] = (p1[1] + p2[1]) / 2.f;
        sph[2] = (p1[2] + p2[2]) / 2.f;
        sph[3] = sqrtf(best) / 2.f;
    }
    __syncthreads();
    for(uint32_t base = 0; base < n; base += SPHERE_BLOCK)
    {
        const uint32_t i = base + tid;
        float px = 0.f, py = 0.f, pz = 0.f;
        if(i < n)
        {
            px = pts[3 * (size_t)i];
            py = pts[3 * (size_t)i + 1];
            pz = pts[3 * (size_t)i + 2];
        }
        uint32_t start = 0;
        while(true)
        {
            const float cx = sph[0], cy = sph[1], cz = sph[2], r = sph[3];
            const float vx = px - cx, vy = py - cy, vz = pz - cz;
            const float d2 = (vx * vx + vy * vy) + vz * vz;
            const bool outside = i < n && tid >= start && d2 > r * r;
            if(tid == 0) first = 0xffffffffu;
            __syncthreads();
            if(outside) atomicMin(&first, tid);
            __syncthreads();
            const uint32_t f = first;
            if(f == 0xffffffffu) break;
            if(tid == f)
            {
                const float d = sqrtf(d2);
                const float k = 0.5f + (r / d) / 2;
                sph[0] = px * (1.f - k) + cx * k;
                sph[1] = py * (1.f - k) + cy * k;
                sph[2] = pz * (1.f - k) + cz * k;
                sph[3] = (r + d) / 2;
            }
            __syncthreads();
            start = f + 1;
        }
        __syncthreads();
    }
    if(tid == 0)
    {
        out4[0] = sph[0];
        out4[1] = sph[1];
        out4[2] = sph[2];
        out4[3] = sph[3];
    }
}

// ---------------------------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------------------------
void launch_tri_prep(hipStream_t s, const float *tris, uint32_t n, float *pre)
{
    if(!n) return;
    k_tri_prep<<<div_up(n, 256), 256, 0, s>>>(tris, n, pre);
}

void launch_fill_u64(hipStream_t s, uint64_t *p, uint32_t n, uint64_t value)
{
    if(!n) return;
    k_fill_u64<<<div_up(n, 256), 256, 0, s>>>(p, n, value);
}

void launch_fill_u8(hipStream_t s, uint8_t *p, uint32_t n, uint8_t value)
{
    if(!n) return;
    k_fill_u8<<<div_up(n, 256), 256, 0, s>>>(p, n, value);
}

static uint32_t brute_chunk(uint32_t n_ray, uint32_t n_tri)
{
    // enough (ray block x triangle chunk) tiles for >= 16k waves, chunks of >= 256 triangles
    uint64_t ray_waves = div_up(n_ray, 64);
    uint64_t want = (16384 + ray_waves - 1) / ray_waves;
    uint64_t max_chunks = div_up(n_tri, 256);
    uint64_t chunks = want < 1 ? 1 : (want > max_chunks ? max_chunks : want);
    if(chunks > 65535) chunks = 65535;
    return div_up(n_tri, chunks);
}

void brute_closest(vkrt_ctx *ctx, const float *pre, uint32_t n_tri, const float4 *rays, uint32_t n_ray,
                   uint64_t *keys, vkrt_hit *hits)
{
    hipStream_t s = ctx->stream;
    launch_fill_u64(s, keys, n_ray, ~0ull);
    if(n_tri)
    {
        uint32_t chunk = brute_chunk(n_ray, n_tri);
        dim3 grid(div_up(n_ray, 256), div_up(n_tri, chunk));
        k_brute_closest<<<grid, 256, 0, s>>>(rays, n_ray, pre, n_tri, chunk, keys);
    }
    k_brute_finalize<<<div_up(n_ray, 256), 256, 0, s>>>(rays, n_ray, pre, keys, hits);
}

void brute_any(vkrt_ctx *ctx, const float *pre, uint32_t n_tri, const float4 *rays, uint32_t n_ray, uint8_t *occ)
{
    hipStream_t s = ctx->stream;
    launch_fill_u8(s, occ, n_ray, 0);
    if(!n_tri) return;
    uint32_t chunk = brute_chunk(n_ray, n_tri);
    dim3 grid(div_up(n_ray, 256), div_up(n_tri, chunk));
    k_brute_any<<<grid, 256, 0, s>>>(rays, n_ray, pre, n_tri, chunk, occ);
}

// ---------------------------------------------------------------------------------------------
// allocation-free brute force: its scratch (the (v0, e1, e2) triangles, then the 64-bit hit keys) is
// a file-scope device array, so a call allocates nothing. Each GPU holds its own copy of the array;
// the host reaches it only through the address hipGetSymbolAddress returns on that device (the
// array's name on the host is a stand-in, not a device address).
// ---------------------------------------------------------------------------------------------
constexpr size_t NOALLOC_TRI_BYTES = (size_t)VKRT_NOALLOC_MAX_TRI * 9 * sizeof(float);
constexpr size_t NOALLOC_BYTES = NOALLOC_TRI_BYTES + (size_t)VKRT_NOALLOC_MAX_RAY * sizeof(uint64_t);
static_assert(NOALLOC_TRI_BYTES % 256 == 0, "the key area must stay 256-B aligned");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];

//! resolves `scratch` on the context's device (current when this runs, from vkrt_create), once per device
void noalloc_scratch_init(vkrt_ctx *ctx)
{
    if(ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) return;// the entry point then reports it
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if(!g_scratch_ptr[ctx->device])
    {
        void *p = nullptr;
        VKRT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        VKRT_REQUIRE(p, "hipGetSymbolAddress returned no address for the allocation-free scratch");
        g_scratch_ptr[ctx->device] = p;
    }
    ctx->noalloc_scratch = g_scratch_ptr[ctx->device];
}

}// namespace vkrt

// =============================================================================================
// C ABI
// =============================================================================================
using namespace vkrt;

template<class Out, class K>
static int pairwise(vkrt_ctx *ctx, const float *prims, size_t floats_per, const vkrt_ray *rays, uint32_t n, Out *out,
                    int flags, K kernel)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(n == 0 || (prims && rays && out), "null array argument");
        if(n == 0) return;
        const float *dp = input(ctx, 0, prims, n * floats_per, flags);
        const vkrt_ray *dr = input(ctx, 1, rays, n, flags);
        Out *dout = output(ctx, 2, out, n, flags);
        kernel<<<div_up(n, 256), 256, 0, ctx->stream>>>(dp, reinterpret_cast<const float4 *>(dr), n, dout);
        VKRT_HIP(hipGetLastError());
        finish_output(ctx, 2, out, n, flags);
        finish_call(ctx, flags);
    });
}

extern "C" {

int vkrt_intersect_triangles(vkrt_ctx *ctx, const float *tris, const vkrt_ray *rays, uint32_t n,
                             vkrt_tri_result *out, int flags)
{
    return pairwise(ctx, tris, 9, rays, n, out, flags, k_pair_tri);
}

int vkrt_intersect_aabbs(vkrt_ctx *ctx, const float *boxes, const vkrt_ray *rays, uint32_t n, vkrt_result *out,
                         int flags)
{
    return pairwise(ctx, boxes, 6, rays, n, out, flags, k_pair_aabb);
}

int vkrt_intersect_obbs(vkrt_ctx *ctx, const float *obbs, const vkrt_ray *rays, uint32_t n, vkrt_result *out,
                        int flags)
{
    return pairwise(ctx, obbs, 15, rays, n, out, flags, k_pair_obb);
}

int vkrt_intersect_spheres(vkrt_ctx *ctx, const float *spheres, const vkrt_ray *rays, uint32_t n,
                           vkrt_result *out, int flags)
{
    return pairwise(ctx, spheres, 4, rays, n, out, flags, k_pair_sphere);
}

int vkrt_intersect_planes(vkrt_ctx *ctx, const float *planes, const vkrt_ray *rays, uint32_t n, vkrt_result *out,
                          int flags)
{
    return pairwise(ctx, planes, 4, rays, n, out, flags, k_pair_plane);
}

int vkrt_tri_box_overlap(vkrt_ctx *ctx, const float *tris, const float *boxes, uint32_t n, uint32_t *out, int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(n == 0 || (tris && boxes && out), "null array argument");
        if(n == 0) return;
        const float *dt = input(ctx, 0, tris, (size_t)n * 9, flags);
        const float *db = input(ctx, 1, boxes, (size_t)n * 6, flags);
        uint32_t *dout = output(ctx, 2, out, n, flags);
        k_pair_tri_box<<<div_up(n, 256), 256, 0, ctx->stream>>>(dt, db, n, dout);
        VKRT_HIP(hipGetLastError());
        finish_output(ctx, 2, out, n, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_tri_tri_overlap(vkrt_ctx *ctx, const float *tris_a, const float *tris_b, uint32_t n, uint32_t *out,
                         int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(n == 0 || (tris_a && tris_b && out), "null array argument");
        if(n == 0) return;
        const float *da = input(ctx, 0, tris_a, (size_t)n * 9, flags);
        const float *db = input(ctx, 1, tris_b, (size_t)n * 9, flags);
        uint32_t *dout = output(ctx, 2, out, n, flags);
        k_pair_tri_tri<<<div_up(n, 256), 256, 0, ctx->stream>>>(da, db, n, dout);
        VKRT_HIP(hipGetLastError());
        finish_output(ctx, 2, out, n, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_compute_aabb(vkrt_ctx *ctx, const float *points, uint32_t n, float *out6, int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(out6 && (n == 0 || points), "null array argument");
        const float *dp = input(ctx, 0, points, (size_t)n * 3, flags);
        float *dout = output(ctx, 2, out6, 6, flags);
        ctx->scratch[0].reserve(6 * sizeof(uint64_t));
        uint64_t *keys = ctx->scratch[0].as<uint64_t>();
        launch_fill_u64(ctx->stream, keys, 3, ~0ull);
        launch_fill_u64(ctx->stream, keys + 3, 3, 0ull);
        if(n)
        {
            uint32_t blocks = div_up(n, 256);
            if(blocks > 4 * (uint32_t)ctx->cu_count) blocks = 4 * ctx->cu_count;
            k_aabb_reduce<<<blocks, 256, 0, ctx->stream>>>(dp, n, keys);
        }
        k_aabb_finalize<<<1, 64, 0, ctx->stream>>>(dp, keys, dout);
        VKRT_HIP(hipGetLastError());
        finish_output(ctx, 2, out6, 6, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_compute_bounding_sphere(vkrt_ctx *ctx, const float *points, uint32_t n, float *out4, int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(out4 && (n == 0 || points), "null array argument");
        float *dout = output(ctx, 2, out4, 4, flags);
        if(n == 0)
        {
            // the reference reads vertices[0] of an empty array; here the sphere is all zeros
            const float zero[4] = {0.f, 0.f, 0.f, 0.f};
            VKRT_HIP(hipMemcpyAsync(dout, zero, sizeof zero, hipMemcpyHostToDevice, ctx->stream));
            VKRT_HIP(hipStreamSynchronize(ctx->stream));
        }
        else
        {
            const float *dp = input(ctx, 0, points, (size_t)n * 3, flags);
            ctx->scratch[0].reserve(6 * sizeof(uint64_t));
            uint64_t *keys = ctx->scratch[0].as<uint64_t>();
            launch_fill_u64(ctx->stream, keys, 3, ~0ull);
            launch_fill_u64(ctx->stream, keys + 3, 3, 0ull);
            uint32_t blocks = div_up(n, 256);
            if(blocks > 4 * (uint32_t)ctx->cu_count) blocks = 4 * ctx->cu_count;
            k_aabb_reduce<<<blocks, 256, 0, ctx->stream>>>(dp, n, keys);
            k_sphere_grow<<<1, SPHERE_BLOCK, 0, ctx->stream>>>(dp, n, keys, dout);
            VKRT_HIP(hipGetLastError());
        }
        finish_output(ctx, 2, out4, 4, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_intersect_brute_closest(vkrt_ctx *ctx, const float *tris, uint32_t n_tri, const vkrt_ray *rays,
                                 uint32_t n_ray, vkrt_hit *out, int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE((n_tri == 0 || tris) && (n_ray == 0 || (rays && out)), "null array argument");
        if(n_ray == 0) return;
        const float *dt = input(ctx, 0, tris, (size_t)n_tri * 9, flags);
        const vkrt_ray *dr = input(ctx, 1, rays, n_ray, flags);
        vkrt_hit *dh = output(ctx, 2, out, n_ray, flags);
        ctx->scratch[0].reserve((size_t)n_tri * 9 * sizeof(float));
        ctx->scratch[1].reserve((size_t)n_ray * sizeof(uint64_t));
        float *pre = ctx->scratch[0].as<float>();
        launch_tri_prep(ctx->stream, dt, n_tri, pre);
        brute_closest(ctx, pre, n_tri, reinterpret_cast<const float4 *>(dr), n_ray, ctx->scratch[1].as<uint64_t>(),
                      dh);
        VKRT_HIP(hipGetLastError());
        finish_output(ctx, 2, out, n_ray, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_intersect_brute_closest_noalloc(vkrt_ctx *ctx, const float *tris, uint32_t n_tri, const vkrt_ray *rays,
                                         uint32_t n_ray, vkrt_hit *out, int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(flags & VKRT_DEVICE_PTRS,
                     "vkrt_intersect_brute_closest_noalloc takes device pointers only (VKRT_DEVICE_PTRS)");
        VKRT_REQUIRE((n_tri == 0 || tris) && (n_ray == 0 || (rays && out)), "null array argument");
        VKRT_REQUIRE(n_tri <= VKRT_NOALLOC_MAX_TRI,
                     "vkrt_intersect_brute_closest_noalloc: n_tri exceeds VKRT_NOALLOC_MAX_TRI (262144)");
        VKRT_REQUIRE(n_ray <= VKRT_NOALLOC_MAX_RAY,
                     "vkrt_intersect_brute_closest_noalloc: n_ray exceeds VKRT_NOALLOC_MAX_RAY (786432)");
        VKRT_REQUIRE(ctx->noalloc_scratch, "no allocation-free scratch on this device");
        if(n_ray == 0) return;
        unsigned char *base = static_cast<unsigned char *>(ctx->noalloc_scratch);
        float *pre = reinterpret_cast<float *>(base);
        uint64_t *keys = reinterpret_cast<uint64_t *>(base + NOALLOC_TRI_BYTES);
        // every word brute_closest reads is written first in the same call: the triangles by the prep
        // kernel, the keys by a fill kernel, so nothing left by an earlier call is read
        launch_tri_prep(ctx->stream, tris, n_tri, pre);
        brute_closest(ctx, pre, n_tri, reinterpret_cast<const float4 *>(rays), n_ray, keys, out);
        VKRT_HIP(hipGetLastError());
        finish_call(ctx, flags);
    });
}

int vkrt_intersect_brute_any(vkrt_ctx *ctx, const float *tris, uint32_t n_tri, const vkrt_ray *rays,
                             uint32_t n_ray, uint8_t *occluded, int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE((n_tri == 0 || tris) && (n_ray == 0 || (rays && occluded)), "null array argument");
        if(n_ray == 0) return;
        const float *dt = input(ctx, 0, tris, (size_t)n_tri * 9, flags);
        const vkrt_ray *dr = input(ctx, 1, rays, n_ray, flags);
        uint8_t *docc = output(ctx, 2, occluded, n_ray, flags);
        ctx->scratch[0].reserve((size_t)n_tri * 9 * sizeof(float));
        float *pre = ctx->scratch[0].as<float>();
        launch_tri_prep(ctx->stream, dt, n_tri, pre);
        brute_any(ctx, pre, n_tri, reinterpret_cast<const float4 *>(dr), n_ray, docc);
        VKRT_HIP(hipGetLastError());
        finish_output(ctx, 2, occluded, n_ray, flags);
        finish_call(ctx, flags);
    });
}

}// extern "C"
