// traverse.hpp - device-side BVH traversal shared by the trace, batched and path-tracer kernels.
//
// One ray per lane, front-to-back descent of the BVH2 (both children's boxes live in the parent, so a
// node visit is four 16-B loads and two slab tests). The per-lane stack keeps its top LDS_DEPTH entries
// in LDS, slot-major with a stride of the block size (lane i of a wave touches bank i: conflict-free),
// and only deeper entries in a private array, which the compiler places in scratch memory.
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
    return r;
}

__device__ __forceinline__ ray_setup setup_ray(const float4 *__restrict__ rays, uint32_t i)
{
    return make_ray_setup(rays[2 * (size_t)i], rays[2 * (size_t)i + 1]);
}

//! Conservative ray/box interval. Each slab value t = RN(RN(b - o) * RN(1/d)) is within 3 eps
//! (eps = 2^-24) relative of (b - o)/d, so growing the interval by T_PAD |t| (T_PAD = 2^-16) keeps the
//! exact interval inside it; boxes were grown by BOX_PAD relative at build time to absorb the rounding
//! of the reference test's hit point.
__device__ __forceinline__ void slab_interval(const ray_setup &r, float lox, float hix, float loy, float hiy, float loz,
                                              float hiz, float &tn, float &tf)
{
    float x0 = (lox - r.o.x) * r.inv.x, x1 = (hix - r.o.x) * r.inv.x;
    float y0 = (loy - r.o.y) * r.inv.y, y1 = (hiy - r.o.y) * r.inv.y;
    float z0 = (loz - r.o.z) * r.inv.z, z1 = (hiz - r.o.z) * r.inv.z;
    float n_ = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1));
    float f_ = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
    tn = n_ - fabsf(n_) * T_PAD;
    tf = f_ + fabsf(f_) * T_PAD;
}

__device__ __forceinline__ void leaf_decode(int ref, uint32_t &first, uint32_t &count)
{
    uint32_t v = static_cast<uint32_t>(~ref);
    first = v >> 3;
    count = (v & 7u) + 1;
}

//! closest (ANY = false) or any (ANY = true) hit with tmin < t < tmax; ties on t go to the lowest
//! original triangle index. `lds` points at this lane's column of a [LDS_DEPTH][BLOCK] int array.
//! With `cull_back`, triangles hit from behind are skipped unless their record is flagged two-sided
//! (bit 0 of the fourth word of the triangle's last float4).
template<bool ANY, int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ bool traverse(const vkrt_bvh_node *__restrict__ nodes, const float4 *__restrict__ tris,
                                         int32_t root, const ray_setup &r, int32_t *lds, float &best_t,
                                         float &best_u, float &best_v, uint32_t &best_prim, bool cull_back = false)
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
                const float4 *tr = tris + 3 * (size_t)(first + k);
                float4 p0 = tr[0], p1 = tr[1], p2 = tr[2];
                float t, u, v;
                const bool cull = cull_back && !(__float_as_uint(p2.w) & TRI_TWO_SIDED);
                if(vierkant::rt::mt_fast(r.o, r.d, vec3{p0.x, p0.y, p0.z}, vec3{p0.w, p1.x, p1.y},
                                         vec3{p1.z, p1.w, p2.x}, t, u, v, cull) &&
                   t > r.tmin)
                {
                    const uint32_t prim = __float_as_uint(p2.y);
                    if(t < best_t || (t == best_t && best_prim != VKRT_INVALID_PRIM && prim < best_prim))
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
        ref = (LDS_DEPTH > 0 && sp < LDS_DEPTH) ? lds[sp * BLOCK] : spill[sp - LDS_DEPTH];
    }
    return best_prim != VKRT_INVALID_PRIM;
}

}// namespace vkrt
