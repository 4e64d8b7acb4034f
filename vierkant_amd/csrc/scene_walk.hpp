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

//! closest (or, ANY, first) hit of world ray (ra, rb) in a two-level scene, in one loop over both levels:
//! top-level and mesh nodes share the 4-wide format, so a lane visits a node of whichever level it is in and
//! the wave stays converged across instance boundaries. Entering an instance leaf switches the node array,
//! triangles and ray to that instance's mesh and object space and remembers the stack depth; when the walk
//! pops back to that depth the mesh is done and the world ray is set up again. Top-level leaves hold one
//! instance each. Ties: smallest t, then the lowest instance, then the lowest triangle.
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
    ray_setup r = make_ray_setup(ra, rb);
    const vkrt_bvh4_node *nodes = tlas;
    const float4 *tris = nullptr;
    uint32_t inst = VKRT_INVALID_PRIM;
    int base = -1;// stack depth at which the current mesh walk began (-1: at the top level)
    int sp = 0;
    int32_t ref = troot;
    while(true)
    {
        if(ref >= 0)
        {
            int32_t c[4];
            const int cnt = visit4(nodes, ref, r, best_t, c);
            if(cnt)
            {
                for(int j = cnt - 1; j >= 1; --j) st.put(sp++, c[j]);
                ref = c[0];
                continue;
            }
        }
        else if(base < 0)
        {
            uint32_t first, count;
            leaf_decode(ref, first, count);
            inst = order[first];
            const inst_dev I = insts[inst];
            const blas_dev B = blases[I.blas];
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
            uint32_t bf, bc;
            leaf_decode(ref, bf, bc);
            bool found = false;
            for(uint32_t k = 0; k < bc; ++k)
            {
                const size_t tb = 48 * (size_t)(bf + k);
                const float4 p0 = load_global_f4(tris, tb), p1 = load_global_f4(tris, tb + 16),
                             p2 = load_global_f4(tris, tb + 32);
                float t, u, v;
                const bool cull = cull_back && !(__float_as_uint(p2.w) & TRI_TWO_SIDED);
                if(vierkant::rt::mt_fast(r.o, r.d, vec3{p0.x, p0.y, p0.z}, vec3{p0.w, p1.x, p1.y},
                                         vec3{p1.z, p1.w, p2.x}, t, u, v, cull) &&
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
                    }
                }
            }
            if(ANY && found) return;
        }
        if(sp == base)
        {
            // the mesh walk is done: back to the top level
            r = make_ray_setup(ra, rb);
            nodes = tlas;
            base = -1;
        }
        if(sp == 0) return;
        --sp;
        ref = st.get(sp);
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
