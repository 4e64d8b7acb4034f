// traverse.hpp - device-side BVH traversal shared by the trace, batched and path-tracer kernels.
//
// One ray per lane, front-to-back descent. traverse() walks the BVH2 (both children's boxes live in the
// parent: a node visit is four 16-B loads and two slab tests); traverse4() walks the 4-wide collapse
// (seven 16-B loads and four slab tests per visit, half the levels), pushing the hit children farthest
// first. Stack entries are child refs only: storing entry distances for pop-time culling doubled the LDS
// stack and lost more to occupancy than the culling saved (3.79 vs 3.90 Grays/s, heightfield scene).
// The per-lane stack keeps its top LDS_DEPTH entries in LDS, slot-major with a stride of the block size
// (lane i of a wave touches bank i: conflict-free), and only deeper entries in a private array, which the
// compiler places in scratch memory.
#pragma once

#include "bvh_types.hpp"
#include "kernels.hpp"
#include "vierkant/rt_math.hpp"

namespace vkrt
{

using vierkant::rt::vec3;

struct ray_setup
{
    vec3 o, d, inv;
    float tmin, tmax;
    vec3 cn, cf;// slab offsets of each axis's near plane (met first) and far plane (see make_ray_setup)
    // byte offsets in a 4-wide node of each axis's near-plane float4 (lo when the direction component is
    // >= 0, else hi); the far plane's float4 is at the offset ^ 16
    uint32_t nx, ny, nz;
};

__device__ __forceinline__ ray_setup make_ray_setup(float4 a, float4 b)
{
    ray_setup r;
    r.o = vec3{a.x, a.y, a.z};
    r.tmin = a.w;
    r.d = vec3{b.x, b.y, b.z};
    r.tmax = b.w;
    // a zero direction component becomes +-1e-20: that axis then gives t = 0 on a slab plane through the
    // origin and |t| >= 1e20 |b - o| elsewhere, i.e. "inside the slab for all t" or "never"
    auto safe = [](float x) { return fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x; };
    r.inv = vec3{1.0f / safe(r.d.x), 1.0f / safe(r.d.y), 1.0f / safe(r.d.z)};
    // a plane b gives t = fmaf(b, inv, c): c = -o inv shifted by A = 2^-20 |o inv| toward smaller t for the
    // plane the ray meets first on that axis (the near plane: lo when inv >= 0, else hi) and toward larger t
    // for the far one. fl(o inv) and the rounding of c each miss o inv by at most 2^-24 |o inv|, far inside A,
    // so the shifted value bounds (b - o) inv from the conservative side on every axis; the relative error
    // left (inv's and the fma's rounding, 2 x 2^-24 |t|) is the one the interval pad T_PAD covers.
    // (an infinite o inv takes no shift: both planes then sit at the same infinite t, which is exact)
    auto offsets = [](float o, float inv, float &cn, float &cf) {
        const float p = o * inv, a = fabsf(p) < __builtin_huge_valf() ? fabsf(p) * 0x1p-20f : 0.0f;
        cn = -p - a;
        cf = -p + a;
    };
    offsets(r.o.x, r.inv.x, r.cn.x, r.cf.x);
    offsets(r.o.y, r.inv.y, r.cn.y, r.cf.y);
    offsets(r.o.z, r.inv.z, r.cn.z, r.cf.z);
    r.nx = r.inv.x >= 0.0f ? 0u : 16u;
    r.ny = r.inv.y >= 0.0f ? 32u : 48u;
    r.nz = r.inv.z >= 0.0f ? 64u : 80u;
    return r;
}

__device__ __forceinline__ ray_setup setup_ray(const float4 *__restrict__ rays, uint32_t i)
{
    return make_ray_setup(rays[2 * (size_t)i], rays[2 * (size_t)i + 1]);
}

//! Conservative ray/box interval. Each slab value t = fmaf(b, inv, c) is (b - o)/d within 2 eps (eps =
//! 2^-24) relative, moved outward by the axis's offset shift (make_ray_setup) which covers the absolute
//! error of o inv; growing the interval by T_PAD |t| (T_PAD = 2^-16) then keeps the exact interval inside it.
//! Boxes were grown by BOX_PAD relative at build time to absorb the rounding of the reference test's hit
//! point.
__device__ __forceinline__ void slab_interval(const ray_setup &r, float lox, float hix, float loy, float hiy, float loz,
                                              float hiz, float &tn, float &tf)
{
    // (the binary-tree walk: each axis's lo plane takes the near offset where inv >= 0, the far one elsewhere)
    const bool px = r.inv.x >= 0.0f, py = r.inv.y >= 0.0f, pz = r.inv.z >= 0.0f;
    float x0 = fmaf(lox, r.inv.x, px ? r.cn.x : r.cf.x), x1 = fmaf(hix, r.inv.x, px ? r.cf.x : r.cn.x);
    float y0 = fmaf(loy, r.inv.y, py ? r.cn.y : r.cf.y), y1 = fmaf(hiy, r.inv.y, py ? r.cf.y : r.cn.y);
    float z0 = fmaf(loz, r.inv.z, pz ? r.cn.z : r.cf.z), z1 = fmaf(hiz, r.inv.z, pz ? r.cf.z : r.cn.z);
    float n_ = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1));
    float f_ = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
    // one rounding per end (a fused multiply-add): at most half an ulp against the 2^-16 |t| of the pad
    tn = fmaf(-fabsf(n_), T_PAD, n_);
    tf = fmaf(fabsf(f_), T_PAD, f_);
}

static_assert(sizeof(vkrt_bvh4_node) == 128, "4-wide node layout: lo/hi float4 pairs at 0/16, 32/48, 64/80");

//! the padded intervals of a 4-wide node's four children and its child refs. Each axis's near-plane float4
//! is loaded from the offset the ray chose at setup and the far-plane float4 from that offset ^ 16, so a
//! child costs two fmas per axis and one max/min per end: for a valid box (lo <= hi) the near plane's
//! value is the smaller of the two on every axis, the same value slab_interval's per-axis min picks. For a
//! box with lo > hi (the empty box of non-finite triangles) the interval comes out empty. Node offsets are
//! 32-bit: the build keeps the 4-wide layout only while the nodes fit 4 GiB.
//! BVH nodes and triangle records always live in device global memory; saying so lets a pointer that a kernel
//! read from memory (a batch item, a scene record) take global loads with a scalar base and a 32-bit lane
//! offset instead of flat loads with a 64-bit address per lane (two VGPRs per load in flight)
template<class T>
using global_ptr = const __attribute__((address_space(1))) T *;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 load_global_f4(const void *base, size_t byte_offset)
{
    const f32x4 v = *(global_ptr<f32x4>)((global_ptr<char>)base + byte_offset);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 load_global_i4(const void *base, size_t byte_offset)
{
    const i32x4 v = *(global_ptr<i32x4>)((global_ptr<char>)base + byte_offset);
    return make_int4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void node4_intervals(const vkrt_bvh4_node *__restrict__ nodes, int32_t ref,
                                                const ray_setup &r, float tn[4], float tf[4], int4 &child)
{
    const uint32_t nb = static_cast<uint32_t>(ref) * static_cast<uint32_t>(sizeof(vkrt_bvh4_node));
    const uint32_t ax = nb + r.nx, ay = nb + r.ny, az = nb + r.nz;
    const float4 nx = load_global_f4(nodes, ax), fx = load_global_f4(nodes, ax ^ 16u);
    const float4 ny = load_global_f4(nodes, ay), fy = load_global_f4(nodes, ay ^ 16u);
    const float4 nz = load_global_f4(nodes, az), fz = load_global_f4(nodes, az ^ 16u);
    child = load_global_i4(nodes, nb + 96u);
    const float n0[4] = {nx.x, nx.y, nx.z, nx.w}, f0[4] = {fx.x, fx.y, fx.z, fx.w};
    const float n1[4] = {ny.x, ny.y, ny.z, ny.w}, f1[4] = {fy.x, fy.y, fy.z, fy.w};
    const float n2[4] = {nz.x, nz.y, nz.z, nz.w}, f2[4] = {fz.x, fz.y, fz.z, fz.w};
#pragma unroll
    for(int j = 0; j < 4; ++j)
    {
        const float n_ = fmaxf(fmaxf(fmaf(n0[j], r.inv.x, r.cn.x), fmaf(n1[j], r.inv.y, r.cn.y)),
                               fmaf(n2[j], r.inv.z, r.cn.z));
        const float f_ = fminf(fminf(fmaf(f0[j], r.inv.x, r.cf.x), fmaf(f1[j], r.inv.y, r.cf.y)),
                               fmaf(f2[j], r.inv.z, r.cf.z));
        tn[j] = fmaf(-fabsf(n_), T_PAD, n_);
        tf[j] = fmaf(fabsf(f_), T_PAD, f_);
    }
}

__device__ __forceinline__ void leaf_decode(int ref, uint32_t &first, uint32_t &count)
{
    uint32_t v = static_cast<uint32_t>(~ref);
    first = v >> 3;
    count = (v & 7u) + 1;
}

//! the candidate filter of a traversal: a candidate that would become the closest hit and whose record is
//! flagged TRI_ALPHA_MASK is kept only if filter(record's last float4, u, v) says so (the any-hit stage);
//! keep_all, the default, compiles the test out
struct keep_all
{
    static constexpr bool active = false;
    __device__ bool operator()(const float4 &, float, float) const { return true; }
};

template<class F>
__device__ __forceinline__ bool filtered_out(const F &filter, const float4 &rec, float u, float v)
{
    if constexpr(F::active)
        return (__float_as_uint(rec.w) & TRI_ALPHA_MASK) && !filter(rec, u, v);
    else
        return false;
}

//! stack entry i of a lane: the LDS column for the top LDS_DEPTH entries, the private spill array past them.
//! The LDS slot is read unconditionally (at a clamped index) and the spill entry only past the LDS depth, so
//! the common pop is an LDS read.
template<int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ int32_t stack_get(const int32_t *lds, const int32_t *spill, int i)
{
    if constexpr(LDS_DEPTH == 0)
        return spill[i];
    else
    {
        int32_t v = lds[(i < LDS_DEPTH ? i : LDS_DEPTH - 1) * BLOCK];
        // (volatile keeps the two loads apart: merged, they become one load through a generic pointer)
        if(i >= LDS_DEPTH) v = *static_cast<const volatile int32_t *>(spill + (i - LDS_DEPTH));
        return v;
    }
}

//! closest (ANY = false) or any (ANY = true) hit with tmin < t < tmax; ties on t go to the lowest
//! original triangle index. `lds` points at this lane's column of a [LDS_DEPTH][BLOCK] int array.
//! With `cull_back`, triangles hit from behind are skipped unless their record is flagged two-sided
//! (bit 0 of the fourth word of the triangle's last float4).
template<bool ANY, int BLOCK, int LDS_DEPTH, class F = keep_all>
__device__ __forceinline__ bool traverse(const vkrt_bvh_node *__restrict__ nodes, const float4 *__restrict__ tris,
                                         int32_t root, const ray_setup &r, int32_t *lds, float &best_t,
                                         float &best_u, float &best_v, uint32_t &best_prim, bool cull_back = false,
                                         const F &filter = F{})
{
    constexpr int SPILL = STACK_SIZE - LDS_DEPTH > 0 ? STACK_SIZE - LDS_DEPTH : 1;
    int32_t spill[SPILL];
    int sp = 0;
    int32_t ref = root;
    best_t = r.tmax;
    best_prim = VKRT_INVALID_PRIM;
    while(true)
    {
        if(ref >= 0)
        {
            const vkrt_bvh_node &nd = nodes[ref];
            float4 a = nd.a, b = nd.b, c = nd.c;
            int4 d = nd.d;
            float tn0, tf0, tn1, tf1;
            slab_interval(r, a.x, a.y, a.z, a.w, c.x, c.y, tn0, tf0);
            slab_interval(r, b.x, b.y, b.z, b.w, c.z, c.w, tn1, tf1);
            const bool h0 = tn0 <= tf0 && tf0 >= r.tmin && tn0 <= best_t;
            const bool h1 = tn1 <= tf1 && tf1 >= r.tmin && tn1 <= best_t;
            if(h0 && h1)
            {
                const bool first0 = tn0 <= tn1;
                const int32_t far = first0 ? d.y : d.x;
                if(LDS_DEPTH > 0 && sp < LDS_DEPTH) lds[sp * BLOCK] = far;
                else spill[sp - LDS_DEPTH] = far;
                ++sp;
                ref = first0 ? d.x : d.y;
                continue;
            }
            if(h0 || h1)
            {
                ref = h0 ? d.x : d.y;
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
                float4 p0 = load_global_f4(tris, tb), p1 = load_global_f4(tris, tb + 16),
                       p2 = load_global_f4(tris, tb + 32);
                float t, u, v;
                const bool cull = cull_back && !(__float_as_uint(p2.w) & TRI_TWO_SIDED);
                if(vierkant::rt::mt_fast(r.o, r.d, vec3{p0.x, p0.y, p0.z}, vec3{p0.w, p1.x, p1.y},
                                         vec3{p1.z, p1.w, p2.x}, t, u, v, cull) &&
                   t > r.tmin)
                {
                    const uint32_t prim = __float_as_uint(p2.y);
                    if((t < best_t || (t == best_t && best_prim != VKRT_INVALID_PRIM && prim < best_prim)) &&
                       !filtered_out(filter, p2, u, v))
                    {
                        best_t = t;
                        best_u = u;
                        best_v = v;
                        best_prim = prim;
                        if(ANY) return true;
                    }
                }
            }
        }
        if(sp == 0) break;
        --sp;
        ref = stack_get<BLOCK, LDS_DEPTH>(lds, spill, sp);
    }
    return best_prim != VKRT_INVALID_PRIM;
}

//! one leaf's triangles (shared by both traversals); returns true when ANY and a hit was found
template<bool ANY, class F = keep_all>
__device__ __forceinline__ bool test_leaf(const float4 *__restrict__ tris, int32_t ref, const ray_setup &r,
                                          float &best_t, float &best_u, float &best_v, uint32_t &best_prim,
                                          bool cull_back, const F &filter = F{})
{
    uint32_t first, count;
    leaf_decode(ref, first, count);
    for(uint32_t k = 0; k < count; ++k)
    {
        // (a 64-bit offset: triangle records may pass 4 GiB, where a 4-wide node array may not)
        const size_t tb = 48 * (size_t)(first + k);
        float4 p0 = load_global_f4(tris, tb), p1 = load_global_f4(tris, tb + 16), p2 = load_global_f4(tris, tb + 32);
        float t, u, v;
        const bool cull = cull_back && !(__float_as_uint(p2.w) & TRI_TWO_SIDED);
        if(vierkant::rt::mt_fast(r.o, r.d, vec3{p0.x, p0.y, p0.z}, vec3{p0.w, p1.x, p1.y}, vec3{p1.z, p1.w, p2.x}, t,
                                 u, v, cull) &&
           t > r.tmin)
        {
            const uint32_t prim = __float_as_uint(p2.y);
            if((t < best_t || (t == best_t && best_prim != VKRT_INVALID_PRIM && prim < best_prim)) &&
               !filtered_out(filter, p2, u, v))
            {
                best_t = t;
                best_u = u;
                best_v = v;
                best_prim = prim;
                if(ANY) return true;
            }
        }
    }
    return false;
}

//! sort key of a 4-wide node's child: +inf when the ray misses it, else its entry distance clamped below
//! +inf (so a hit child never sorts with the missed ones)
constexpr float KEY_MISSED = __builtin_huge_valf();
//! a child that meets the interval has an entry below +inf (an entry n = +inf pads to fma(-inf, T_PAD, inf) =
//! NaN, which fails every comparison of child_hit), so its own entry is its key
__device__ __forceinline__ float sort_key(bool hit, float k) { return hit ? k : KEY_MISSED; }

//! does the ray's padded interval [tn, tf] of child c meet [tmin, best_t]? The terms are joined with '&',
//! not '&&': a short-circuit puts the far-side reduction in a branch of its own per child, where the
//! compiler no longer knows its operands are fma results and re-quiets each one before fminf/fmaxf
__device__ __forceinline__ bool child_hit(int32_t c, float tn, float tf, float tmin, float best_t)
{
    return (c != EMPTY_CHILD) & (tn <= tf) & (tf >= tmin) & (tn <= best_t);
}

//! branchless compare-exchange of (key, ref) pairs: the smaller key first
__device__ __forceinline__ void compare_exchange(float &ka, int32_t &ca, float &kb, int32_t &cb)
{
    const bool sw = kb < ka;
    const float lo = sw ? kb : ka, hi = sw ? ka : kb;
    const int32_t cl = sw ? cb : ca, ch = sw ? ca : cb;
    ka = lo;
    kb = hi;
    ca = cl;
    cb = ch;
}

//! per-ray traversal state of the 4-wide walk (the stack lives in LDS + a private spill array)
struct trav_state
{
    ray_setup r;
    float best_t, best_u, best_v;
    uint32_t best_prim;
    int32_t ref;
    int sp;
};

__device__ __forceinline__ void trav_init(trav_state &s, const ray_setup &r, int32_t root)
{
    s.r = r;
    s.best_t = r.tmax;
    s.best_u = 0.f;
    s.best_v = 0.f;
    s.best_prim = VKRT_INVALID_PRIM;
    s.ref = root;
    s.sp = 0;
}

//! one visit of the 4-wide walk: a node (children whose padded interval meets [tmin, best_t] are entered
//! nearest first, the others pushed farthest first) or a leaf, then a pop. Returns true when the ray is
//! finished (stack empty, or the first hit when ANY). `lds` is this lane's column of a [LDS_DEPTH][BLOCK]
//! array; `spill` holds entries past LDS_DEPTH. A child is entered under the same condition as in
//! traverse(), and the closest-hit rule does not depend on the visiting order, so the result is the same.
template<bool ANY, int BLOCK, int LDS_DEPTH, class F = keep_all>
__device__ __forceinline__ bool step4(const vkrt_bvh4_node *__restrict__ nodes, const float4 *__restrict__ tris,
                                      trav_state &s, int32_t *lds, int32_t *spill, bool cull_back,
                                      const F &filter = F{})
{
    const ray_setup &r = s.r;
    if(s.ref >= 0)
    {
        float tn[4], tf[4];
        int4 cr;
        node4_intervals(nodes, s.ref, r, tn, tf, cr);
        const float k0 = tn[0], k1 = tn[1], k2 = tn[2], k3 = tn[3];
        bool v0 = child_hit(cr.x, k0, tf[0], r.tmin, s.best_t);
        bool v1 = child_hit(cr.y, k1, tf[1], r.tmin, s.best_t);
        bool v2 = child_hit(cr.z, k2, tf[2], r.tmin, s.best_t);
        bool v3 = child_hit(cr.w, k3, tf[3], r.tmin, s.best_t);
        int32_t c0 = cr.x, c1 = cr.y, c2 = cr.z, c3 = cr.w;
        // sort nearest first with the missed children last (sort keys: +inf for them, the entry distance
        // clamped below +inf for the others): five branchless compare-exchanges of (key, ref)
        float q0 = sort_key(v0, k0), q1 = sort_key(v1, k1), q2 = sort_key(v2, k2), q3 = sort_key(v3, k3);
        compare_exchange(q0, c0, q1, c1);
        compare_exchange(q2, c2, q3, c3);
        compare_exchange(q0, c0, q2, c2);
        compare_exchange(q1, c1, q3, c3);
        compare_exchange(q1, c1, q2, c2);
        if(q0 < KEY_MISSED)
        {
            auto push = [&](int32_t c) {
                if(LDS_DEPTH > 0 && s.sp < LDS_DEPTH) lds[s.sp * BLOCK] = c;
                else spill[s.sp - LDS_DEPTH] = c;
                ++s.sp;
            };
            if(q3 < KEY_MISSED) push(c3);
            if(q2 < KEY_MISSED) push(c2);
            if(q1 < KEY_MISSED) push(c1);
            s.ref = c0;
            return false;
        }
    }
    else if(test_leaf<ANY, F>(tris, s.ref, r, s.best_t, s.best_u, s.best_v, s.best_prim, cull_back, filter))
        return true;
    if(s.sp == 0) return true;
    --s.sp;
    s.ref = stack_get<BLOCK, LDS_DEPTH>(lds, spill, s.sp);
    return false;
}

//! children of 4-wide node `ref` whose padded interval meets [lo, best_t] (lo: the ray's tmin unless given),
//! nearest first; returns the count
__device__ __forceinline__ int visit4(const vkrt_bvh4_node *__restrict__ nodes, int32_t ref, const ray_setup &r,
                                      float best_t, int32_t c[4], float lo)
{
    float k[4], tf[4];
    int4 cr;
    node4_intervals(nodes, ref, r, k, tf, cr);
    bool v[4];
    v[0] = child_hit(cr.x, k[0], tf[0], lo, best_t);
    v[1] = child_hit(cr.y, k[1], tf[1], lo, best_t);
    v[2] = child_hit(cr.z, k[2], tf[2], lo, best_t);
    v[3] = child_hit(cr.w, k[3], tf[3], lo, best_t);
    c[0] = cr.x;
    c[1] = cr.y;
    c[2] = cr.z;
    c[3] = cr.w;
    float q[4] = {sort_key(v[0], k[0]), sort_key(v[1], k[1]), sort_key(v[2], k[2]), sort_key(v[3], k[3])};
    compare_exchange(q[0], c[0], q[1], c[1]);
    compare_exchange(q[2], c[2], q[3], c[3]);
    compare_exchange(q[0], c[0], q[2], c[2]);
    compare_exchange(q[1], c[1], q[3], c[3]);
    compare_exchange(q[1], c[1], q[2], c[2]);
    return (q[0] < KEY_MISSED ? 1 : 0) + (q[1] < KEY_MISSED ? 1 : 0) + (q[2] < KEY_MISSED ? 1 : 0) +
           (q[3] < KEY_MISSED ? 1 : 0);
}

__device__ __forceinline__ int visit4(const vkrt_bvh4_node *__restrict__ nodes, int32_t ref, const ray_setup &r,
                                      float best_t, int32_t c[4])
{
    return visit4(nodes, ref, r, best_t, c, r.tmin);
}

template<int BLOCK, int LDS_DEPTH>
struct lane_stack
{
    int32_t *lds;
    int32_t *spill;
    __device__ __forceinline__ void put(int i, int32_t v)
    {
        if(LDS_DEPTH > 0 && i < LDS_DEPTH) lds[i * BLOCK] = v;
        else spill[i - LDS_DEPTH] = v;
    }
    __device__ __forceinline__ int32_t get(int i) const
    {
        return stack_get<BLOCK, LDS_DEPTH>(lds, spill, i);
    }
};

constexpr int spill_entries(int lds_depth) { return STACK_SIZE - lds_depth > 0 ? STACK_SIZE - lds_depth : 1; }

//! the whole 4-wide walk for one ray (see step4)
template<bool ANY, int BLOCK, int LDS_DEPTH, class F = keep_all>
__device__ __forceinline__ bool traverse4(const vkrt_bvh4_node *__restrict__ nodes, const float4 *__restrict__ tris,
                                          int32_t root, const ray_setup &r, int32_t *lds, float &best_t,
                                          float &best_u, float &best_v, uint32_t &best_prim, bool cull_back = false,
                                          const F &filter = F{})
{
    int32_t spill[spill_entries(LDS_DEPTH)];
    trav_state s;
    trav_init(s, r, root);
    while(!step4<ANY, BLOCK, LDS_DEPTH, F>(nodes, tris, s, lds, spill, cull_back, filter)) {}
    best_t = s.best_t;
    best_u = s.best_u;
    best_v = s.best_v;
    best_prim = s.best_prim;
    return best_prim != VKRT_INVALID_PRIM;
}

//! the 4-wide walk with leaves postponed ("while-while" with speculative traversal, Aila and Laine 2009): a
//! lane that meets a leaf parks it and keeps descending until every lane of the wave still walking holds a
//! parked leaf; then the wave tests its leaves together, so node visits and triangle tests are not
//! interleaved across lanes. The same children are entered under the same condition (with a best_t that can
//! lag behind: more nodes, never fewer) and the closest-hit rule does not depend on the visiting order, so
//! the result is traverse4's.
template<bool ANY, int BLOCK, int LDS_DEPTH, class F = keep_all>
__device__ __forceinline__ bool traverse4_postponed(const vkrt_bvh4_node *__restrict__ nodes,
                                                    const float4 *__restrict__ tris, int32_t root, const ray_setup &r,
                                                    int32_t *lds, float &best_t, float &best_u, float &best_v,
                                                    uint32_t &best_prim, bool cull_back = false,
                                                    const F &filter = F{})
{
    constexpr int32_t NONE = EMPTY_CHILD, DONE = EMPTY_CHILD - 1;// neither is a node index or a leaf ref
    int32_t spill[spill_entries(LDS_DEPTH)];
    lane_stack<BLOCK, LDS_DEPTH> st{lds, spill};
    int sp = 0;
    best_t = r.tmax;
    best_u = best_v = 0.f;
    best_prim = VKRT_INVALID_PRIM;
    auto pop = [&]() -> int32_t { return sp == 0 ? DONE : st.get(--sp); };
    int32_t node = root, leaf = NONE;
    if(node < 0)
    {
        leaf = node;
        node = DONE;
    }
    while(true)
    {
        while(node >= 0 && node != DONE)
        {
            float tn[4], tf[4];
            int4 cr;
            node4_intervals(nodes, node, r, tn, tf, cr);
            const float k0 = tn[0], k1 = tn[1], k2 = tn[2], k3 = tn[3];
            bool v0 = child_hit(cr.x, k0, tf[0], r.tmin, best_t);
            bool v1 = child_hit(cr.y, k1, tf[1], r.tmin, best_t);
            bool v2 = child_hit(cr.z, k2, tf[2], r.tmin, best_t);
            bool v3 = child_hit(cr.w, k3, tf[3], r.tmin, best_t);
            int32_t c0 = cr.x, c1 = cr.y, c2 = cr.z, c3 = cr.w;
            float q0 = sort_key(v0, k0), q1 = sort_key(v1, k1), q2 = sort_key(v2, k2), q3 = sort_key(v3, k3);
            compare_exchange(q0, c0, q1, c1);
            compare_exchange(q2, c2, q3, c3);
            compare_exchange(q0, c0, q2, c2);
            compare_exchange(q1, c1, q3, c3);
            compare_exchange(q1, c1, q2, c2);
            if(q0 < KEY_MISSED)
            {
                if(q3 < KEY_MISSED) st.put(sp++, c3);
                if(q2 < KEY_MISSED) st.put(sp++, c2);
                if(q1 < KEY_MISSED) st.put(sp++, c1);
                node = c0;
            }
            else
                node = pop();
            if(node < 0 && leaf == NONE)
            {
                leaf = node;
                node = pop();
            }
            if(__all(leaf != NONE)) break;
        }
        while(leaf != NONE)
        {
            if(test_leaf<ANY, F>(tris, leaf, r, best_t, best_u, best_v, best_prim, cull_back, filter)) return true;
            if(node < 0)
            {
                leaf = node;
                node = pop();
            }
            else
                leaf = NONE;
        }
        if(node == DONE) break;
    }
    return best_prim != VKRT_INVALID_PRIM;
}

}// namespace vkrt
