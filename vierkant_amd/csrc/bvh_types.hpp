// bvh_types.hpp - BVH layout shared by the build, the traversal kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "context.hpp"

//! BVH2 node holding both children's boxes (64 B, four 16-B loads)
struct vkrt_bvh_node
{
    float4 a;// child 0: lo.x hi.x lo.y hi.y
    float4 b;// child 1: lo.x hi.x lo.y hi.y
    float4 c;// lo.z hi.z of child 0, lo.z hi.z of child 1
    int4 d;  // child refs: >= 0 internal node, < 0 leaf ~((first << 3) | (count - 1))
};

//! 4-wide node collapsed from the BVH2 (128 B): children's boxes by plane, so one node visit is seven
//! 16-B loads and four slab tests; an unused slot has child ref EMPTY_CHILD (its box is never read)
struct vkrt_bvh4_node
{
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 child;// >= 0 BVH4 node, < 0 leaf (the BVH2 leaf encoding), EMPTY_CHILD unused
    int4 src;  // per slot (BVH2 node << 1) | side holding the child's box, -1 unused (read by a refit only)
};

struct vkrt_bvh
{
    int device = 0;
    uint32_t n_tri = 0, n_vert = 0, n_nodes = 0, max_leaf = 4;
    uint32_t builder = 0;// VKRT_BUILDER_*
    uint32_t ploc_radius = 0;// 0: default
    int32_t root_ref = 0;
    bool empty = true;
    vkrt::device_buffer nodes, tris;
    // the 4-wide collapse used by traversal when its worst-case stack fits (wide == true)
    vkrt::device_buffer nodes4, collapse_q[2], collapse_cnt;
    int32_t root4 = 0;
    uint32_t n_nodes4 = 0, stack4 = 0;
    bool wide = false, allow_wide = true;
    vkrt::device_buffer indices, order, child, range, parent, nbox, tbox, pre, err;
    vkrt_bvh_info info{};
};


namespace vkrt
{

constexpr int32_t MAX_LEAF = 8;
constexpr int STACK_SIZE = 64;      // traversal stack entries per ray; the build rejects deeper trees
constexpr float BOX_PAD = 0x1p-16f; // relative growth of every stored box coordinate
constexpr float T_PAD = 0x1p-16f;   // relative growth of every slab interval
constexpr uint32_t TRI_TWO_SIDED = 1u;// flag bit in a leaf triangle record: never culled as a back face
constexpr uint32_t TRI_ALPHA_MASK = 2u;// flag bit: the record's material is alpha-masked (any-hit test)
constexpr int32_t EMPTY_CHILD = 0x7fffffff;// unused BVH4 child slot (no node index reaches it)

}// namespace vkrt
