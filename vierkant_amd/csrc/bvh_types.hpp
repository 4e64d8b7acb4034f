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

struct vkrt_bvh
{
    int device = 0;
    uint32_t n_tri = 0, n_vert = 0, n_nodes = 0, max_leaf = 4;
    int32_t root_ref = 0;
    bool empty = true;
    vkrt::device_buffer nodes, tris;
    vkrt::device_buffer indices, order, child, range, parent, nbox, tbox, flags, pre, err;
    vkrt_bvh_info info{};
};


namespace vkrt
{

constexpr int32_t MAX_LEAF = 8;
constexpr int STACK_SIZE = 64;      // traversal stack entries per ray; the build rejects deeper trees
constexpr float BOX_PAD = 0x1p-16f; // relative growth of every stored box coordinate
constexpr float T_PAD = 0x1p-16f;   // relative growth of every slab interval
constexpr uint32_t TRI_TWO_SIDED = 1u;// flag bit in a leaf triangle record: never culled as a back face

}// namespace vkrt
