// scene_walk.hpp - device-side walk of a two-level scene (a top-level 4-wide BVH over instances of per-mesh
// 4-wide BVHs), shared by the scene trace kernels and the path tracer.
//
// Replaces the driver's instance traversal behind RayBuilder's top-level structure (create_toplevel,
// src/RayBuilder.cpp:386-781; entry_t::transform / inv_transform, include/vierkant/RayBuilder.hpp:24-38).
// Every instance carries its world-to-object matrix; the object-space ray is a fixed fp32 expression of the
// caller's numbers, so the result is defined exactly (oracle: orc_scene_closest).
#pragma once

#include "traverse.hpp"

namespace vkrt
{

constexpr int SCENE_STACK = 96;  // top-level + BLAS entries per ray; checked at build time

struct inst_dev
{
    float4 w0, w1, w2;// world-to-object rows
    uint32_t blas, custom, pad0, pad1;
};

struct blas_dev
{
    const vkrt_bvh4_node *nodes;
    const float4 *tris;
    int32_t root;
    uint32_t empty;
};

//! object-space setup of world ray (ra, rb) for instance I: each row ((w0 x + w1 y) + w2 z) + w3, the direction
//! keeping its length (so t is shared by both spaces)
__device__ __forceinline__ ray_setup object_ray(const inst_dev &I, const float4 &ra, const float4 &rb)
{
    const float ox = ((I.w0.x * ra.x + I.w0.y * ra.y) + I.w0.z * ra.z) + I.w0.w;
    const float oy = ((I.w1.x * ra.x + I.w1.y * ra.y) + I.w1.z * ra.z) + I.w1.w;
    const float oz = ((I.w2.x * ra.x + I.w2.y * ra.y) + I.w2.z * ra.z) + I.w2.w;
    const float dx = (I.w0.x * rb.x + I.w0.y * rb.y) + I.w0.z * rb.z;
    const float dy = (I.w1.x * rb.x + I.w1.y * rb.y) + I.w1.z * rb.z;
    const float dz = (I.w2.x * rb.x + I.w2.y * rb.y) + I.w2.z * rb.z;
    return make_ray_setup(make_float4(ox, oy, oz, ra.w), make_float4(dx, dy, dz, rb.w));
}

constexpr int SCENE_CAND = 4;     // instances kept per top-level pass of scene_walk

//! one mesh leaf's triangles against the object ray, (t, instance, triangle) order; true when one became the
//! best record
template<bool ANY, class F>
__device__ __forceinline__ bool mesh_leaf(const float4 *__restrict__ tris, int32_t ref, const ray_setup &r,
                                          uint32_t inst, float &best_t, float &best_u, float &best_v,
                                          uint32_t &best_prim, uint32_t &best_inst, bool cull_back, const F &filter)
{
    uint32_t first, count;
    leaf_decode(ref, first, count);
    bool found = false;
    for(uint32_t k = 0; k < count; ++k)
    {
        const size_t tb = 48 * (size_t)(first + k);
        const float4 p0 = load_global_f4(tris, tb), p1 = load_global_f4(tris, tb + 16),
                     p2 = load_global_f4(tris, tb + 32);
        float t, u, v;
        const bool cull = cull_back && !(__float_as_uint(p2.w) & TRI_TWO_SIDED);
        if(vierkant::rt::mt_fast(r.o, r.d, vec3{p0.x, p0.y, p0.z}, vec3{p0.w, p1.x, p1.y}, vec3{p1.z, p1.w, p2.x}, t,
                                 u, v, cull) &&
           t > r.tmin)
        {
            const uint32_t prim = __float_as_uint(p2.y);
            if((t < best_t || (t == best_t && best_prim != VKRT_INVALID_PRIM &&
                               (inst < best_inst || (inst == best_inst && prim < best_prim)))) &&
               !filtered_out(filter, p2, u, v))
            {
                best_t = t;
                best_u = u;
                best_v = v;
                best_prim = prim;
                best_inst = inst;
                found = true;
                if(ANY) return true;
            }
        }
    }
    return found;
}

//! one instance's mesh walked with its object ray, which stays fixed for the whole walk (the loop then
//! carries no ray state from one level to another), the best record carried in and out; the step
//! structure of the flat 4-wide walk. Returns true when ANY found a hit
template<bool ANY, int BLOCK, int LDS_DEPTH, class F>
__device__ __forceinline__ bool mesh_walk(const vkrt_bvh4_node *__restrict__ nodes, const float4 *__restrict__ tris,
                                          int32_t root, const ray_setup &r, lane_stack<BLOCK, LDS_DEPTH> &st,
                                          uint32_t inst, float &best_t, float &best_u, float &best_v,
                                          uint32_t &best_prim, uint32_t &best_inst, bool cull_back, const F &filter)
{
    int sp = 0;
    int32_t ref = root;
    while(true)
    {
        if(ref >= 0)
        {
            float tn[4], tf[4];
            int4 cr;
            node4_intervals(nodes, ref, r, tn, tf, cr);
            const float k0 = tn[0], k1 = tn[1], k2 = tn[2], k3 = tn[3];
            const bool v0 = child_hit(cr.x, k0, tf[0], r.tmin, best_t);
            const bool v1 = child_hit(cr.y, k1, tf[1], r.tmin, best_t);
            const bool v2 = child_hit(cr.z, k2, tf[2], r.tmin, best_t);
            const bool v3 = child_hit(cr.w, k3, tf[3], r.tmin, best_t);
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
                ref = c0;
                continue;
            }
        }
        else if(mesh_leaf<ANY, F>(tris, ref, r, inst, best_t, best_u, best_v, best_prim, best_inst, cull_back,
                                  filter) &&
                ANY)
            return true;
        if(sp == 0) return false;
        ref = st.get(--sp);
    }
}

//! closest (or, ANY, first) hit of world ray (ra, rb) in a two-level scene, in passes of two phases. The top
//! level is walked first (nearest child first, boxes beyond the best hit skipped) and the SCENE_CAND instance
//! leaves with the nearest padded box entries are kept in (entry, instance) order; then the kept instances
//! are walked nearest first, each with its object ray fixed for the whole mesh walk, until an entry lies
//! beyond the best hit. A full list may have dropped farther instances, so another pass takes the instances
//! after the last one walked in (entry, instance) order. The entries are the same in every pass (the same
//! arithmetic on the same ray), the walked set grows strictly, and every instance whose box the ray enters
//! before the best hit is walked once. Top-level leaves hold one instance each. Ties: smallest t, then the
//! lowest instance, then the lowest triangle, whatever order the instances are walked in.
//! cull_back drops candidates seen from behind (object-space determinant below zero) unless their record is
//! flagged TRI_TWO_SIDED; a candidate that would become the closest and whose record is flagged
//! TRI_ALPHA_MASK is kept only if filter(record's last float4, u, v) says so (keep_all: no any-hit stage)
template<bool ANY, int BLOCK, int LDS_DEPTH, class F = keep_all>
__device__ __forceinline__ void scene_walk(const vkrt_bvh4_node *__restrict__ tlas, int32_t troot,
                                           const uint32_t *__restrict__ order, const inst_dev *__restrict__ insts,
                                           const blas_dev *__restrict__ blases, const float4 &ra, const float4 &rb,
                                           int32_t *lds, float &best_t, float &best_u, float &best_v,
                                           uint32_t &best_prim, uint32_t &best_inst, bool cull_back = false,
                                           const F &filter = F{})
{
    int32_t spill[SCENE_STACK - LDS_DEPTH];
    lane_stack<BLOCK, LDS_DEPTH> st{lds, spill};
    bool first_pass = true;
    float lb_k = 0.f;   // the last instance walked, in (entry, instance) order
    uint32_t lb_i = 0u;
    while(true)
    {
        float ck[SCENE_CAND];
        uint32_t ci[SCENE_CAND];
#pragma unroll
        for(int j = 0; j < SCENE_CAND; ++j)
        {
            ck[j] = __int_as_float(0x7f800000);
            ci[j] = 0xffffffffu;
        }
        int cc = 0;
        bool more = false;
        auto keep = [&](float k, uint32_t inst) {
            if(!first_pass && (k < lb_k || (k == lb_k && inst <= lb_i))) return;
            if(cc == SCENE_CAND && (k > ck[SCENE_CAND - 1] || (k == ck[SCENE_CAND - 1] && inst > ci[SCENE_CAND - 1])))
            {
                more = true;
                return;
            }
            if(cc == SCENE_CAND) more = true;// the farthest kept one is pushed out
            float kk = k;
            uint32_t ii = inst;
#pragma unroll
            for(int j = 0; j < SCENE_CAND; ++j)
            {
                const bool less = kk < ck[j] || (kk == ck[j] && ii < ci[j]);
                const float tk = ck[j];
                const uint32_t ti = ci[j];
                ck[j] = less ? kk : tk;
                ci[j] = less ? ii : ti;
                kk = less ? tk : kk;
                ii = less ? ti : ii;
            }
            cc = cc < SCENE_CAND ? cc + 1 : SCENE_CAND;
        };
        // the world ray's setup is rebuilt for each pass rather than kept live through the mesh walks (the
        // empty asm keeps the compiler from hoisting it out of the loop)
        float4 a = ra, b = rb;
        asm volatile("" : "+v"(a.x), "+v"(a.y), "+v"(a.z), "+v"(a.w), "+v"(b.x), "+v"(b.y), "+v"(b.z), "+v"(b.w));
        const ray_setup rw = make_ray_setup(a, b);
        if(troot < 0)
        {
            // a single instance: no box to enter, walked whatever the ray
            uint32_t first, count;
            leaf_decode(troot, first, count);
            keep(-__int_as_float(0x7f800000), order[first]);
        }
        else
        {
            int sp = 0;
            int32_t ref = troot;
            while(true)
            {
                // boxes beyond the best hit, or beyond the farthest kept entry once the list is full, are skipped
                const float bound = cc == SCENE_CAND ? fminf(best_t, ck[SCENE_CAND - 1]) : best_t;
                float tn[4], tf[4];
                int4 cr;
                node4_intervals(tlas, ref, rw, tn, tf, cr);
                const bool v0 = child_hit(cr.x, tn[0], tf[0], rw.tmin, bound);
                const bool v1 = child_hit(cr.y, tn[1], tf[1], rw.tmin, bound);
                const bool v2 = child_hit(cr.z, tn[2], tf[2], rw.tmin, bound);
                const bool v3 = child_hit(cr.w, tn[3], tf[3], rw.tmin, bound);
                int32_t c0 = cr.x, c1 = cr.y, c2 = cr.z, c3 = cr.w;
                float q0 = sort_key(v0, tn[0]), q1 = sort_key(v1, tn[1]), q2 = sort_key(v2, tn[2]),
                      q3 = sort_key(v3, tn[3]);
                compare_exchange(q0, c0, q1, c1);
                compare_exchange(q2, c2, q3, c3);
                compare_exchange(q0, c0, q2, c2);
                compare_exchange(q1, c1, q3, c3);
                compare_exchange(q1, c1, q2, c2);
                // instance leaves are kept with their entry; of the inner children the nearest is visited next
                // and the others pushed farthest first
                int32_t next = 0;
                bool have = false;
                auto take = [&](float q, int32_t c) {
                    if(!(q < KEY_MISSED)) return;
                    if(c < 0)
                    {
                        uint32_t first, count;
                        leaf_decode(c, first, count);
                        keep(q, order[first]);
                    }
                    else
                    {
                        if(have) st.put(sp++, next);
                        next = c;
                        have = true;
                    }
                };
                take(q3, c3);
                take(q2, c2);
                take(q1, c1);
                take(q0, c0);
                if(have)
                {
                    ref = next;
                    continue;
                }
                if(sp == 0) break;
                ref = st.get(--sp);
            }
        }
        // a full list may have skipped subtrees by its farthest entry: their instances are the next pass's
        if(cc == SCENE_CAND) more = true;
        bool done = !more;
        for(int j = 0; j < SCENE_CAND; ++j)
        {
            if(j >= cc) break;
            if(ck[j] > best_t)
            {
                done = true;
                break;
            }
            const uint32_t inst = ci[j];
            lb_k = ck[j];
            lb_i = inst;
            const inst_dev I = insts[inst];
            const blas_dev B = blases[I.blas];
            if(B.empty) continue;
            const ray_setup ro = object_ray(I, ra, rb);
            if(mesh_walk<ANY, BLOCK, LDS_DEPTH, F>(B.nodes, B.tris, B.root, ro, st, inst, best_t, best_u, best_v,
                                                   best_prim, best_inst, cull_back, filter))
                return;
        }
        if(done) return;
        first_pass = false;
    }
}

//! device pointers of a two-level scene's top level, instances and meshes (kept by the scene object)
struct scene_view
{
    const vkrt_bvh4_node *tlas;
    int32_t root;
    uint32_t empty;
    const uint32_t *order;
    const inst_dev *insts;
    const blas_dev *blases;
};

scene_view scene_device_view(const vkrt_scene *sc);

}// namespace vkrt
