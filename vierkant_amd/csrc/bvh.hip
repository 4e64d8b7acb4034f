// bvh.hip - mesh -> BVH build (LBVH) and closest/any-hit traversal for gfx950.
//
// The reference hands triangles to the Vulkan driver (RayBuilder::create_mesh_structures,
// src/RayBuilder.cpp:98-303, vkCmdBuildAccelerationStructuresKHR :274) and traces with TraceRay
// (shaders/slang/raypipeline.slang:182); both are opaque. This file is the MI355X replacement.
//
// Build (all on the GPU, one stream):
//   k_gather    indexed mesh -> per-triangle (v0, e1, e2), exact AABB, centroid-bounds atomics
//   k_morton    30-bit Morton code of the AABB center in the centroid bounds
//   radix sort  (code, triangle) pairs, project-written (radix_sort.hip)
//   k_karras    binary radix tree over the sorted codes (Karras 2012), ties broken by position; or, with
//               VKRT_BUILDER_PLOC, agglomerative clustering over the sorted leaves (k_ploc_*)
//   k_box_*     node boxes as unions over each node's leaf run (segment prefix/suffix unions and a sparse
//               table of segment unions)
//   k_layout    BVH2 nodes holding both children's boxes (64 B), subtrees of <= max_leaf_size
//               triangles collapsed into leaves over the contiguous sorted range
//   k_reorder   triangles in leaf order as 48-B records (v0, e1, e2, original index)
//
// Traversal: one ray per lane, front-to-back descent with a per-lane stack. Box tests are padded so
// that a box is never culled when it holds a triangle the reference test accepts (see slab_interval),
// and equal-t hits resolve to the lowest original triangle index, so results are identical to the
// brute-force loop over intersect(Triangle, Ray).
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "context.hpp"
#include "kernels.hpp"
#include "bvh_types.hpp"
#include "radix_sort.hpp"
#include "vierkant/rt_math.hpp"

using namespace vierkant::rt;

namespace vkrt
{

__device__ __forceinline__ vec3 ld3(const float *p) { return vec3{p[0], p[1], p[2]}; }

__device__ __forceinline__ bool finite3(const vec3 &v)
{
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

__device__ __forceinline__ void atomic_min_f(uint32_t *p, float v) { atomicMin(p, ordered_bits(v + 0.0f)); }
__device__ __forceinline__ void atomic_max_f(uint32_t *p, float v) { atomicMax(p, ordered_bits(v + 0.0f)); }

__device__ __forceinline__ float from_ordered(uint32_t b)
{
    uint32_t r = (b & 0x80000000u) ? (b & 0x7fffffffu) : ~b;
    return __uint_as_float(r);
}

// ---------------------------------------------------------------------------------------------
// build kernels
// ---------------------------------------------------------------------------------------------
__global__ void k_gather(const uint8_t *__restrict__ pos, uint32_t stride, const uint32_t *__restrict__ idx,
                         uint32_t n_tri, uint32_t n_vert, float *__restrict__ pre, float *__restrict__ box,
                         uint32_t *__restrict__ cbounds, uint32_t *__restrict__ err)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    // the centroid bounds are reduced across the wave first and reach memory as one atomic per bound and wave
    // (one per triangle serialized 1.8 M triangles' updates on six words: about 1.9 ms); lanes past the end
    // stay in the wave for the reduction
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    if(j < n_tri)
    {
        uint32_t i0 = idx[3 * (size_t)j], i1 = idx[3 * (size_t)j + 1], i2 = idx[3 * (size_t)j + 2];
        if(i0 >= n_vert || i1 >= n_vert || i2 >= n_vert)
        {
            atomicOr(err, 1u);
            i0 = i1 = i2 = 0;
        }
        vec3 v0 = ld3(reinterpret_cast<const float *>(pos + (size_t)i0 * stride));
        vec3 v1 = ld3(reinterpret_cast<const float *>(pos + (size_t)i1 * stride));
        vec3 v2 = ld3(reinterpret_cast<const float *>(pos + (size_t)i2 * stride));
        vec3 e1 = v1 - v0, e2 = v2 - v0;
        float *p = pre + 9 * (size_t)j;
        p[0] = v0.x, p[1] = v0.y, p[2] = v0.z, p[3] = e1.x, p[4] = e1.y, p[5] = e1.z, p[6] = e2.x, p[7] = e2.y,
        p[8] = e2.z;
        float *b = box + 6 * (size_t)j;
        if(finite3(v0) && finite3(v1) && finite3(v2))
        {
            vec3 lo{fminf(fminf(v0.x, v1.x), v2.x), fminf(fminf(v0.y, v1.y), v2.y), fminf(fminf(v0.z, v1.z), v2.z)};
            vec3 hi{fmaxf(fmaxf(v0.x, v1.x), v2.x), fmaxf(fmaxf(v0.y, v1.y), v2.y), fmaxf(fmaxf(v0.z, v1.z), v2.z)};
            b[0] = lo.x, b[1] = lo.y, b[2] = lo.z, b[3] = hi.x, b[4] = hi.y, b[5] = hi.z;
            cmin[0] = cmax[0] = 0.5f * lo.x + 0.5f * hi.x;
            cmin[1] = cmax[1] = 0.5f * lo.y + 0.5f * hi.y;
            cmin[2] = cmax[2] = 0.5f * lo.z + 0.5f * hi.z;
        }
        else
        {
            // a triangle with a non-finite vertex never yields a finite t (the reference test then produces
            // NaN), so it gets an empty box and is never visited
            b[0] = b[1] = b[2] = FLT_MAX;
            b[3] = b[4] = b[5] = -FLT_MAX;
        }
    }
    for(int off = 32; off > 0; off >>= 1)
        for(int k = 0; k < 3; ++k)
        {
            cmin[k] = fminf(cmin[k], __shfl_xor(cmin[k], off));
            cmax[k] = fmaxf(cmax[k], __shfl_xor(cmax[k], off));
        }
    if((threadIdx.x & 63u) == 0 && cmin[0] <= cmax[0])
    {
        atomic_min_f(&cbounds[0], cmin[0]), atomic_min_f(&cbounds[1], cmin[1]), atomic_min_f(&cbounds[2], cmin[2]);
        atomic_max_f(&cbounds[3], cmax[0]), atomic_max_f(&cbounds[4], cmax[1]), atomic_max_f(&cbounds[5], cmax[2]);
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ void k_morton(const float *__restrict__ box, uint32_t n, const uint32_t *__restrict__ cbounds,
                         uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if(j >= n) return;
    const float *b = box + 6 * (size_t)j;
    uint32_t code = 0;
    if(b[0] <= b[3])
    {
        uint32_t q[3];
        for(int a = 0; a < 3; ++a)
        {
            float lo = from_ordered(cbounds[a]), hi = from_ordered(cbounds[3 + a]);
            float c = 0.5f * b[a] + 0.5f * b[3 + a];
            float ext = hi - lo;
            float f = ext > 0.f ? (c - lo) / ext : 0.5f;
            f = fminf(fmaxf(f, 0.f), 1.f);
            q[a] = min(1023u, static_cast<uint32_t>(f * 1024.f));
        }
        code = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    }
    keys[j] = code;
    vals[j] = j;
}

__device__ __forceinline__ int delta(const uint32_t *__restrict__ keys, int n, int i, int j)
{
    if(j < 0 || j >= n) return -1;
    uint32_t a = keys[i], b = keys[j];
    if(a == b) return 32 + __clz(static_cast<uint32_t>(i ^ j));
    return __clz(a ^ b);
}

//! Karras 2012 binary radix tree. child refs: >= 0 internal node, < 0 leaf at sorted position ~ref.
__global__ void k_karras(const uint32_t *__restrict__ keys, int n, int2 *__restrict__ child, uint2 *__restrict__ range,
                         int32_t *__restrict__ parent_int, int32_t *__restrict__ parent_leaf)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while(delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for(int t = lmax / 2; t >= 1; t /= 2)
        if(delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    for(int div = 2;; div *= 2)
    {
        int t = (l + div - 1) / div;
        if(delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if(t <= 1) break;
    }
    const int gamma = i + s * d + min(d, 0);
    const int first = min(i, j), last = max(i, j);
    int left = (first == gamma) ? ~gamma : gamma;
    int right = (last == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    child[i] = make_int2(left, right);
    range[i] = make_uint2(first, last);
    if(left < 0) parent_leaf[~left] = i;
    else parent_int[left] = i;
    if(right < 0) parent_leaf[~right] = i;
    else parent_int[right] = i;
    if(i == 0) parent_int[0] = -1;
}

__global__ void k_sorted_boxes(const uint32_t *__restrict__ order, uint32_t n, const float *__restrict__ box,
                               float *__restrict__ tbox)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if(j >= n) return;
    const float *s = box + 6 * (size_t)order[j];
    float *d = tbox + 6 * (size_t)j;
    for(int a = 0; a < 6; ++a) d[a] = s[a];
}

__device__ __forceinline__ void load_box(const float *p, float b[6])
{
    for(int a = 0; a < 6; ++a) b[a] = p[a];
}

//! Node boxes without a bottom-up hand-off. Every internal node covers a contiguous run [first, last] of the
//! leaf order (`range`), so its box is the union of those leaves' boxes. The leaves are cut into segments of
//! BOX_SEG: k_box_scan writes, per leaf, the union from its segment's start (pre) and to its segment's end
//! (suf), and each segment's union (level 0 of a sparse table, levels 1.. by k_box_sparse: level k holds the
//! union of 2^k segments from each start). A run that crosses segments is suf[first], pre[last] and at most two
//! overlapping table entries; a run inside one segment is pre or suf when it touches the segment's start or
//! end, else a loop over its (at most BOX_SEG - 2) leaves. Every pass reads what earlier kernels wrote, so no
//! box crosses between workgroups, or XCDs, inside a kernel. A min or max is exact whatever the grouping, so
//! the boxes are the ones the pairwise union up the tree gives; only leaves' boxes are combined (the padding
//! past the last leaf is never read into a union).
constexpr uint32_t BOX_SEG = 64;

__device__ __forceinline__ void box_union(float a[6], const float b[6])
{
    for(int k = 0; k < 3; ++k) a[k] = fminf(a[k], b[k]);
    for(int k = 3; k < 6; ++k) a[k] = fmaxf(a[k], b[k]);
}

//! one segment of BOX_SEG leaves per workgroup (one wave): inclusive prefix and suffix unions in LDS
__global__ void __launch_bounds__(BOX_SEG) k_box_scan(uint32_t n, const float *__restrict__ tbox,
                                                      float *__restrict__ pre, float *__restrict__ suf,
                                                      float *__restrict__ seg)
{
    __shared__ float sp[6][BOX_SEG], ss[6][BOX_SEG];
    const uint32_t t = threadIdx.x, j = blockIdx.x * BOX_SEG + t;
    const uint32_t cnt = min(BOX_SEG, n - blockIdx.x * BOX_SEG);// leaves in this segment
    float a[6], b[6];
    if(t < cnt) load_box(tbox + 6 * (size_t)j, a);
    else
        for(int k = 0; k < 6; ++k) a[k] = 0.f;// past the last leaf: never combined into a leaf's union
    for(int k = 0; k < 6; ++k) b[k] = a[k];
    for(uint32_t off = 1; off < BOX_SEG; off <<= 1)
    {
        for(int k = 0; k < 6; ++k)
        {
            sp[k][t] = a[k];
            ss[k][t] = b[k];
        }
        __syncthreads();
        if(t >= off)
        {
            float o[6];
            for(int k = 0; k < 6; ++k) o[k] = sp[k][t - off];
            box_union(a, o);
        }
        if(t + off < cnt)
        {
            float o[6];
            for(int k = 0; k < 6; ++k) o[k] = ss[k][t + off];
            box_union(b, o);
        }
        __syncthreads();
    }
    if(t < cnt)
    {
        for(int k = 0; k < 6; ++k)
        {
            pre[6 * (size_t)j + k] = a[k];
            suf[6 * (size_t)j + k] = b[k];
        }
    }
    if(t == 0)
        for(int k = 0; k < 6; ++k) seg[6 * (size_t)blockIdx.x + k] = b[k];// leaf 0's suffix: the segment
}

//! sparse-table level k from level k - 1: entry i = union of entries i and i + 2^(k-1), for i + 2^k <= nseg
__global__ void k_box_sparse(uint32_t nseg, uint32_t k, float *__restrict__ table)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t half = 1u << (k - 1);
    if((size_t)i + 2 * half > nseg) return;
    const float *lo = table + 6 * ((size_t)(k - 1) * nseg + i);
    const float *hi = table + 6 * ((size_t)(k - 1) * nseg + i + half);
    float a[6], b[6];
    load_box(lo, a);
    load_box(hi, b);
    box_union(a, b);
    for(int c = 0; c < 6; ++c) table[6 * ((size_t)k * nseg + i) + c] = a[c];
}

__global__ void k_box_nodes(uint32_t ni, uint32_t n, uint32_t nseg, const uint2 *__restrict__ range,
                            const float *__restrict__ tbox, const float *__restrict__ pre,
                            const float *__restrict__ suf, const float *__restrict__ table, float *__restrict__ nbox)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= ni) return;
    const uint2 r = range[p];
    const uint32_t first = r.x, last = r.y, sf = first / BOX_SEG, sl = last / BOX_SEG;
    float o[6], x[6];
    if(sf == sl)
    {
        const bool at_end = last % BOX_SEG == BOX_SEG - 1 || last == n - 1;
        if(first % BOX_SEG == 0) load_box(pre + 6 * (size_t)last, o);
        else if(at_end) load_box(suf + 6 * (size_t)first, o);
        else
        {
            load_box(tbox + 6 * (size_t)first, o);
            for(uint32_t j = first + 1; j <= last; ++j)
            {
                load_box(tbox + 6 * (size_t)j, x);
                box_union(o, x);
            }
        }
    }
    else
    {
        load_box(suf + 6 * (size_t)first, o);
        load_box(pre + 6 * (size_t)last, x);
        box_union(o, x);
        if(sl > sf + 1)
        {
            const uint32_t lo = sf + 1, count = sl - 1 - lo + 1;
            const uint32_t k = 31u - __builtin_clz(count);
            load_box(table + 6 * ((size_t)k * nseg + lo), x);
            box_union(o, x);
            load_box(table + 6 * ((size_t)k * nseg + (sl - 1) - (1u << k) + 1), x);
            box_union(o, x);
        }
    }
    for(int c = 0; c < 6; ++c) nbox[6 * (size_t)p + c] = o[c];
}

//! the sparse table's depth for nseg segments: levels 0 .. floor(log2(nseg))
static uint32_t box_levels(uint32_t nseg)
{
    uint32_t l = 1;
    while((2ull << (l - 1)) <= nseg) ++l;
    return l;
}

__device__ __forceinline__ float pad_lo(float x) { return x - fabsf(x) * BOX_PAD; }
__device__ __forceinline__ float pad_hi(float x) { return x + fabsf(x) * BOX_PAD; }

__global__ void k_layout(uint32_t n, const int2 *__restrict__ child, const uint2 *__restrict__ range,
                         const float *__restrict__ nbox, const float *__restrict__ tbox, uint32_t max_leaf,
                         vkrt_bvh_node *__restrict__ nodes)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i + 1 >= n) return;
    int2 c = child[i];
    float bx[2][6];
    int ref[2];
    const int cr[2] = {c.x, c.y};
    for(int k = 0; k < 2; ++k)
    {
        uint32_t first, last;
        const float *src;
        if(cr[k] < 0)
        {
            first = last = ~cr[k];
            src = tbox + 6 * (size_t)first;
        }
        else
        {
            uint2 r = range[cr[k]];
            first = r.x, last = r.y;
            src = nbox + 6 * (size_t)cr[k];
        }
        uint32_t count = last - first + 1;
        ref[k] = (cr[k] < 0 || count <= max_leaf) ? ~static_cast<int>((first << 3) | (count - 1)) : cr[k];
        for(int a = 0; a < 3; ++a) bx[k][a] = pad_lo(src[a]);
        for(int a = 3; a < 6; ++a) bx[k][a] = pad_hi(src[a]);
    }
    vkrt_bvh_node nd;
    nd.a = make_float4(bx[0][0], bx[0][3], bx[0][1], bx[0][4]);
    nd.b = make_float4(bx[1][0], bx[1][3], bx[1][1], bx[1][4]);
    nd.c = make_float4(bx[0][2], bx[0][5], bx[1][2], bx[1][5]);
    nd.d = make_int4(ref[0], ref[1], 0, 0);
    nodes[i] = nd;
}

__global__ void k_reorder(const uint32_t *__restrict__ order, uint32_t n, const float *__restrict__ pre,
                          float4 *__restrict__ tris)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if(j >= n) return;
    uint32_t src = order[j];
    const float *p = pre + 9 * (size_t)src;
    tris[3 * (size_t)j + 0] = make_float4(p[0], p[1], p[2], p[3]);
    tris[3 * (size_t)j + 1] = make_float4(p[4], p[5], p[6], p[7]);
    tris[3 * (size_t)j + 2] = make_float4(p[8], __uint_as_float(src), 0.f, 0.f);
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
//! build statistics over the final BVH2 nodes. A node stays internal in the layout when its run holds more
//! than max_leaf triangles (every ancestor's run then does too), so those nodes and the root are the ones a
//! traversal reaches. Each counts its depth up the parent chain (the root is 1), and adds its children's SAH
//! terms (box area; a leaf's times its triangle count, divided by the root's area on the host), its leaf
//! children and their depth. Sums go through a wave reduction and one atomic per wave.
struct bvh_stats_dev
{
    double sah;
    uint32_t leaves, depth;
};

__device__ __forceinline__ double child_area(float lx, float hx, float ly, float hy, float lz, float hz)
{
    const double dx = fmax(0.0, (double)hx - lx), dy = fmax(0.0, (double)hy - ly), dz = fmax(0.0, (double)hz - lz);
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

__global__ void k_stats(uint32_t ni, uint32_t max_leaf, const vkrt_bvh_node *__restrict__ nodes,
                        const uint2 *__restrict__ range, const int32_t *__restrict__ parent_int,
                        bvh_stats_dev *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double sah = 0.0;
    uint32_t leaves = 0, depth = 0;
    if(i < ni)
    {
        const uint2 r = range[i];
        if(i == 0 || r.y - r.x + 1 > max_leaf)
        {
            // (past STACK_SIZE + 1 the walk stops: the host rejects any depth above STACK_SIZE)
            uint32_t d = 1;
            for(int32_t p = static_cast<int32_t>(i); p > 0 && d <= STACK_SIZE + 1; ++d) p = parent_int[p];
            const vkrt_bvh_node nd = nodes[i];
            const int32_t refs[2] = {nd.d.x, nd.d.y};
            const double ar[2] = {child_area(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.c.x, nd.c.y),
                                  child_area(nd.b.x, nd.b.y, nd.b.z, nd.b.w, nd.c.z, nd.c.w)};
            depth = d;
            for(int k = 0; k < 2; ++k)
            {
                if(refs[k] >= 0) sah += ar[k];
                else
                {
                    const uint32_t v = static_cast<uint32_t>(~refs[k]);
                    sah += ((v & 7u) + 1) * ar[k];
                    ++leaves;
                    depth = d + 1;
                }
            }
        }
    }
    for(int off = 32; off > 0; off >>= 1)
    {
        sah += __shfl_xor(sah, off);
        leaves += __shfl_xor(leaves, off);
        depth = max(depth, (uint32_t)__shfl_xor(depth, off));
    }
    if((threadIdx.x & 63u) == 0 && (leaves || depth))
    {
        atomicAdd(&out->sah, sah);
        atomicAdd(&out->leaves, leaves);
        atomicMax(&out->depth, depth);
    }
}

//! the root's box from its two child boxes: rb as (lo, hi) pairs per axis, bounds as lo xyz then hi xyz
static void root_bounds(const vkrt_bvh_node &root, float rb[6], float bounds[6])
{
    const float r[6] = {std::min(root.a.x, root.b.x), std::max(root.a.y, root.b.y), std::min(root.a.z, root.b.z),
                        std::max(root.a.w, root.b.w), std::min(root.c.x, root.c.z), std::max(root.c.y, root.c.w)};
    for(int a = 0; a < 6; ++a) rb[a] = r[a];
    for(int a = 0; a < 3; ++a)
    {
        bounds[a] = r[2 * a];
        bounds[3 + a] = r[2 * a + 1];
    }
}

static void compute_stats(vkrt_bvh *bvh, hipStream_t s)
{
    // depth and SAH cost from the final nodes (build-time statistics only)
    vkrt_bvh_info &info = bvh->info;
    info.depth = 1;
    info.sah_cost = 0.f;
    info.n_leaves = bvh->empty ? 0 : 1;
    if(bvh->empty) return;
    if(bvh->root_ref < 0)
    {
        // a single leaf (at most 8 triangles): the union of its triangles' boxes (tbox, in leaf order),
        // padded like every node box
        const uint32_t v = static_cast<uint32_t>(~bvh->root_ref), first = v >> 3, count = (v & 7u) + 1;
        std::vector<float> box(6 * (size_t)count);
        VKRT_HIP(hipMemcpyAsync(box.data(), bvh->tbox.as<float>() + 6 * (size_t)first, box.size() * 4,
                                hipMemcpyDeviceToHost, s));
        VKRT_HIP(hipStreamSynchronize(s));
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for(uint32_t k = 0; k < count; ++k)
            for(int a = 0; a < 3; ++a)
            {
                lo[a] = std::min(lo[a], box[6 * (size_t)k + a]);
                hi[a] = std::max(hi[a], box[6 * (size_t)k + 3 + a]);
            }
        for(int a = 0; a < 3; ++a)
        {
            info.bounds[a] = lo[a] - std::fabs(lo[a]) * BOX_PAD;
            info.bounds[3 + a] = hi[a] + std::fabs(hi[a]) * BOX_PAD;
        }
        return;
    }
    // the sums on the GPU (a wave reduction and one atomic per wave), the root node and the sums read back
    bvh_stats_dev *dst = reinterpret_cast<bvh_stats_dev *>(bvh->err.as<uint8_t>() + 32);
    VKRT_HIP(hipMemsetAsync(dst, 0, sizeof(bvh_stats_dev), s));
    k_stats<<<div_up(bvh->n_nodes, 256), 256, 0, s>>>(bvh->n_nodes, bvh->max_leaf, bvh->nodes.as<vkrt_bvh_node>(),
                                                      bvh->range.as<uint2>(), bvh->parent.as<int32_t>(), dst);
    VKRT_HIP(hipGetLastError());
    bvh_stats_dev st;
    vkrt_bvh_node root;
    VKRT_HIP(hipMemcpyAsync(&st, dst, sizeof st, hipMemcpyDeviceToHost, s));
    VKRT_HIP(hipMemcpyAsync(&root, bvh->nodes.ptr, sizeof root, hipMemcpyDeviceToHost, s));
    VKRT_HIP(hipStreamSynchronize(s));
    float rb[6];
    root_bounds(root, rb, info.bounds);
    const double dx = std::max(0.0, (double)rb[1] - rb[0]), dy = std::max(0.0, (double)rb[3] - rb[2]),
                 dz = std::max(0.0, (double)rb[5] - rb[4]);
    const double root_area = std::max(1e-30, 2.0 * (dx * dy + dy * dz + dz * dx));
    info.depth = std::max(1u, st.depth);
    info.sah_cost = static_cast<float>(1.0 + st.sah / root_area);// 1: the root's own visit
    info.n_leaves = st.leaves;
    if(info.depth > (uint32_t)STACK_SIZE) throw error(VKRT_E_INTERNAL, "BVH deeper than the traversal stack");
}

static void build_topology_ploc(vkrt_ctx *ctx, vkrt_bvh *bvh);

static void build_topology(vkrt_ctx *ctx, vkrt_bvh *bvh, const uint8_t *pos, uint32_t stride, const uint32_t *idx)
{
    hipStream_t s = ctx->stream;
    const uint32_t n = bvh->n_tri;
    bvh->pre.reserve((size_t)n * 9 * sizeof(float));
    bvh->tbox.reserve((size_t)n * 6 * sizeof(float));
    bvh->err.reserve(64);
    // scratch: per-triangle boxes in input order + centroid bounds + sort keys/values + sort temp
    device_buffer &tmp = ctx->scratch[2];
    const size_t box_bytes = (size_t)n * 6 * sizeof(float);
    const size_t key_bytes = (size_t)n * sizeof(uint32_t);
    tmp.reserve(box_bytes + 64 + key_bytes + radix_sort_temp_bytes(n));
    float *box = tmp.as<float>();
    uint32_t *cb = reinterpret_cast<uint32_t *>(tmp.as<uint8_t>() + box_bytes);
    uint32_t *keys = reinterpret_cast<uint32_t *>(tmp.as<uint8_t>() + box_bytes + 64);
    void *sort_tmp = tmp.as<uint8_t>() + box_bytes + 64 + key_bytes;
    bvh->order.reserve(key_bytes);
    uint32_t *order = bvh->order.as<uint32_t>();

    uint32_t init[8] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u, 0u, 0u};
    VKRT_HIP(hipMemcpyAsync(cb, init, sizeof(init), hipMemcpyHostToDevice, s));
    VKRT_HIP(hipMemsetAsync(bvh->err.ptr, 0, 4, s));
    k_gather<<<div_up(n, 256), 256, 0, s>>>(pos, stride, idx, n, bvh->n_vert, bvh->pre.as<float>(), box, cb,
                                            bvh->err.as<uint32_t>());
    k_morton<<<div_up(n, 256), 256, 0, s>>>(box, n, cb, keys, order);
    radix_sort_pairs(s, keys, order, n, 30, sort_tmp);
    k_sorted_boxes<<<div_up(n, 256), 256, 0, s>>>(order, n, box, bvh->tbox.as<float>());
    uint32_t err_host = 0;
    VKRT_HIP(hipMemcpyAsync(&err_host, bvh->err.ptr, 4, hipMemcpyDeviceToHost, s));
    if(n >= 2)
    {
        const uint32_t ni = n - 1;
        bvh->child.reserve(ni * sizeof(int2));
        bvh->range.reserve(ni * sizeof(uint2));
        bvh->parent.reserve(((size_t)ni + n) * sizeof(int32_t));
        bvh->nbox.reserve((size_t)ni * 6 * sizeof(float));
        int32_t *parent_int = bvh->parent.as<int32_t>(), *parent_leaf = parent_int + ni;
        if(bvh->builder == VKRT_BUILDER_PLOC) build_topology_ploc(ctx, bvh);
        else
            k_karras<<<div_up(ni, 256), 256, 0, s>>>(keys, (int)n, bvh->child.as<int2>(), bvh->range.as<uint2>(),
                                                     parent_int, parent_leaf);
    }
    VKRT_HIP(hipStreamSynchronize(s));
    VKRT_HIP(hipGetLastError());
    if(err_host) throw error(VKRT_E_INVALID_ARG, "triangle index out of range of the vertex count");
}

// ---------------------------------------------------------------------------------------------
// PLOC (Meister & Bittner 2018): agglomerative clustering over the Morton-sorted leaves. Each round every
// cluster finds, within PLOC_RADIUS positions of the sorted list, the cluster whose merged box has the
// smallest surface area; mutual nearest neighbours merge into a new node at the lower position, and the list
// is compacted. Pairs are ordered by (area, lower position, higher position), so the globally smallest pair
// is always mutual and every round merges. Internal nodes are numbered in merge order, then renumbered so the
// root is node 0 and parents precede children; leaves are finally put in depth-first order so that every
// subtree covers a contiguous range, which the leaf collapse of k_layout and the refit rely on.
// ---------------------------------------------------------------------------------------------
constexpr int PLOC_RADIUS = 16;    // default search radius
constexpr int PLOC_MAX_RADIUS = 32;
constexpr int PLOC_BLOCK = 256;
constexpr uint32_t PLOC_MAX_DEPTH = 4096;// bound of the upward walks; deeper trees are rejected after layout

__device__ __forceinline__ float merged_half_area(const float4 &alo, const float4 &ahi, const float4 &blo,
                                                  const float4 &bhi)
{
    const float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
    const float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
    const float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
    return dx * dy + dy * dz + dz * dx;
}

__global__ void k_ploc_init(const float *__restrict__ tbox, uint32_t n, int32_t *__restrict__ cl_id,
                            float4 *__restrict__ lo, float4 *__restrict__ hi, int32_t *__restrict__ parent_int,
                            uint32_t ni)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if(j < ni) parent_int[j] = -1;
    if(j >= n) return;
    const float *b = tbox + 6 * (size_t)j;
    cl_id[j] = ~static_cast<int32_t>(j);
    lo[j] = make_float4(b[0], b[1], b[2], 0.f);
    hi[j] = make_float4(b[3], b[4], b[5], 0.f);
}

//! nearest neighbour of every cluster within the window; ties go to the smallest position, which is the
//! smallest pair in (area, lower, higher) order
__global__ void __launch_bounds__(PLOC_BLOCK) k_ploc_nn(const float4 *__restrict__ lo, const float4 *__restrict__ hi,
                                                        uint32_t nc, int radius, int32_t *__restrict__ nn)
{
    __shared__ float4 slo[PLOC_BLOCK + 2 * PLOC_MAX_RADIUS], shi[PLOC_BLOCK + 2 * PLOC_MAX_RADIUS];
    const int base = static_cast<int>(blockIdx.x * PLOC_BLOCK) - radius;
    for(int k = threadIdx.x; k < PLOC_BLOCK + 2 * radius; k += PLOC_BLOCK)
    {
        const int g = base + k;
        if(g >= 0 && g < static_cast<int>(nc))
        {
            slo[k] = lo[g];
            shi[k] = hi[g];
        }
    }
    __syncthreads();
    const int i = static_cast<int>(blockIdx.x * PLOC_BLOCK + threadIdx.x);
    if(i >= static_cast<int>(nc)) return;
    const int li = threadIdx.x + radius;
    float best = 0.f;
    int bj = -1;
    for(int o = -radius; o <= radius; ++o)
    {
        const int j = i + o;
        if(o == 0 || j < 0 || j >= static_cast<int>(nc)) continue;
        const float d = merged_half_area(slo[li], shi[li], slo[li + o], shi[li + o]);
        if(bj < 0 || d < best)
        {
            best = d;
            bj = j;
        }
    }
    nn[i] = bj;
}

__global__ void k_ploc_flags(const int32_t *__restrict__ nn, uint32_t nc, uint32_t *__restrict__ merge,
                             uint32_t *__restrict__ keep)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if(i >= static_cast<int>(nc)) return;
    const int j = nn[i];
    const bool mutual = nn[j] == i;
    merge[i] = (mutual && i < j) ? 1u : 0u;
    keep[i] = (mutual && i > j) ? 0u : 1u;
}

__device__ __forceinline__ uint32_t leaf_count(const uint32_t *__restrict__ cnt, int32_t ref)
{
    return ref < 0 ? 1u : cnt[ref];
}

__global__ void k_ploc_merge(const int32_t *__restrict__ nn, uint32_t nc, const uint32_t *__restrict__ merge,
                             const uint32_t *__restrict__ moff, const uint32_t *__restrict__ keep,
                             const uint32_t *__restrict__ koff, uint32_t base, const int32_t *__restrict__ cl_id,
                             const float4 *__restrict__ lo, const float4 *__restrict__ hi,
                             int32_t *__restrict__ out_id, float4 *__restrict__ out_lo, float4 *__restrict__ out_hi,
                             int2 *__restrict__ child, uint32_t *__restrict__ cnt, int32_t *__restrict__ parent_int,
                             int32_t *__restrict__ parent_leaf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= nc || !keep[i]) return;
    const uint32_t pos = koff[i];
    if(!merge[i])
    {
        out_id[pos] = cl_id[i];
        out_lo[pos] = lo[i];
        out_hi[pos] = hi[i];
        return;
    }
    const uint32_t j = static_cast<uint32_t>(nn[i]);
    const int32_t id = static_cast<int32_t>(base + moff[i]);
    const int32_t a = cl_id[i], b = cl_id[j];
    child[id] = make_int2(a, b);
    cnt[id] = leaf_count(cnt, a) + leaf_count(cnt, b);
    if(a < 0) parent_leaf[~a] = id;
    else parent_int[a] = id;
    if(b < 0) parent_leaf[~b] = id;
    else parent_int[b] = id;
    const float4 la = lo[i], lb = lo[j], ha = hi[i], hb = hi[j];
    out_id[pos] = id;
    out_lo[pos] = make_float4(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z), 0.f);
    out_hi[pos] = make_float4(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z), 0.f);
}

//! internal node k becomes ni - 1 - k: the last merge (the root) is node 0 and parents precede children
__global__ void k_ploc_renumber(uint32_t ni, uint32_t n, const int2 *__restrict__ child, const uint32_t *__restrict__ cnt,
                                const int32_t *__restrict__ parent_int, const int32_t *__restrict__ parent_leaf,
                                int2 *__restrict__ child2, uint32_t *__restrict__ cnt2, int32_t *__restrict__ parent_int2,
                                int32_t *__restrict__ parent_leaf2)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    auto map = [ni](int32_t r) { return r < 0 ? r : static_cast<int32_t>(ni - 1u) - r; };
    if(k < ni)
    {
        const uint32_t nk = ni - 1u - k;
        const int2 c = child[k];
        child2[nk] = make_int2(map(c.x), map(c.y));
        cnt2[nk] = cnt[k];
        parent_int2[nk] = parent_int[k] < 0 ? -1 : map(parent_int[k]);
    }
    if(k < n) parent_leaf2[k] = map(parent_leaf[k]);
}

//! depth-first position of every leaf and the first position of every internal node's range: the leaf
//! counts of all left siblings on the way up
__global__ void k_ploc_dfs(uint32_t ni, uint32_t n, const int2 *__restrict__ child, const uint32_t *__restrict__ cnt,
                           const int32_t *__restrict__ parent_int, const int32_t *__restrict__ parent_leaf,
                           uint32_t *__restrict__ leaf_pos, uint2 *__restrict__ range)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k < n)
    {
        uint32_t off = 0;
        int32_t cur = ~static_cast<int32_t>(k), node = parent_leaf[k];
        for(uint32_t step = 0; node >= 0 && step < PLOC_MAX_DEPTH; ++step)
        {
            const int2 c = child[node];
            if(c.y == cur) off += leaf_count(cnt, c.x);
            cur = node;
            node = parent_int[node];
        }
        leaf_pos[k] = off;
    }
    if(k < ni)
    {
        uint32_t off = 0;
        int32_t cur = static_cast<int32_t>(k), node = parent_int[k];
        for(uint32_t step = 0; node >= 0 && step < PLOC_MAX_DEPTH; ++step)
        {
            const int2 c = child[node];
            if(c.y == cur) off += leaf_count(cnt, c.x);
            cur = node;
            node = parent_int[node];
        }
        range[k] = make_uint2(off, off + cnt[k] - 1u);
    }
}

//! leaves to depth-first order: the triangle order, sorted boxes, leaf parents and leaf child refs
__global__ void k_ploc_permute(uint32_t ni, uint32_t n, const uint32_t *__restrict__ leaf_pos,
                               const uint32_t *__restrict__ order_in, const float *__restrict__ tbox_in,
                               const int32_t *__restrict__ parent_leaf_in, uint32_t *__restrict__ order_out,
                               float *__restrict__ tbox_out, int32_t *__restrict__ parent_leaf_out,
                               const int2 *__restrict__ child_in, int2 *__restrict__ child_out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k < n)
    {
        const uint32_t d = leaf_pos[k];
        if(d >= n) return;// only a malformed tree gets here; compute_stats rejects it
        order_out[d] = order_in[k];
        for(int a = 0; a < 6; ++a) tbox_out[6 * (size_t)d + a] = tbox_in[6 * (size_t)k + a];
        parent_leaf_out[d] = parent_leaf_in[k];
    }
    if(k < ni)
    {
        const int2 c = child_in[k];
        auto map = [leaf_pos](int32_t r) { return r < 0 ? ~static_cast<int32_t>(leaf_pos[~r]) : r; };
        child_out[k] = make_int2(map(c.x), map(c.y));
    }
}

__global__ void __launch_bounds__(1024) k_scan_local(const uint32_t *__restrict__ in, uint32_t n,
                                                     uint32_t *__restrict__ out, uint32_t *__restrict__ block_sums)
{
    __shared__ uint32_t wsum[16];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    const uint32_t x = i < n ? in[i] : 0u;
    uint32_t sc = x;
    for(int off = 1; off < 64; off <<= 1)
    {
        const uint32_t y = __shfl_up(sc, off);
        if(lane >= (uint32_t)off) sc += y;
    }
    if(lane == 63) wsum[wave] = sc;
    __syncthreads();
    if(threadIdx.x < 16)
    {
        const uint32_t w = wsum[threadIdx.x];
        uint32_t acc = w;
        for(int off = 1; off < 16; off <<= 1)
        {
            const uint32_t y = __shfl_up(acc, off, 16);
            if(threadIdx.x >= (uint32_t)off) acc += y;
        }
        wsum[threadIdx.x] = acc - w;
        if(threadIdx.x == 15) block_sums[blockIdx.x] = acc;
    }
    __syncthreads();
    if(i < n) out[i] = wsum[wave] + sc - x;
}

//! exclusive scan of the per-block sums in one workgroup, 1024 at a time with a running carry
__global__ void __launch_bounds__(1024) k_scan_blocks(uint32_t *__restrict__ data, uint32_t n)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if(threadIdx.x == 0) carry = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for(uint32_t base = 0; base < n; base += 1024)
    {
        const uint32_t i = base + threadIdx.x;
        const uint32_t x = i < n ? data[i] : 0u;
        uint32_t sc = x;
        for(int off = 1; off < 64; off <<= 1)
        {
            const uint32_t y = __shfl_up(sc, off);
            if(lane >= (uint32_t)off) sc += y;
        }
        if(lane == 63) wsum[wave] = sc;
        __syncthreads();
        if(threadIdx.x < 16)
        {
            const uint32_t w = wsum[threadIdx.x];
            uint32_t acc = w;
            for(int off = 1; off < 16; off <<= 1)
            {
                const uint32_t y = __shfl_up(acc, off, 16);
                if(threadIdx.x >= (uint32_t)off) acc += y;
            }
            wsum[threadIdx.x] = acc - w;
        }
        __syncthreads();
        const uint32_t c = carry;
        if(i < n) data[i] = c + wsum[wave] + sc - x;
        __syncthreads();
        if(threadIdx.x == 1023) carry = c + wsum[wave] + sc;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(1024) k_scan_add(const uint32_t *__restrict__ in, uint32_t n,
                                                   uint32_t *__restrict__ out, const uint32_t *__restrict__ block_off,
                                                   uint32_t *__restrict__ total)
{
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    if(i >= n) return;
    const uint32_t v = out[i] + block_off[blockIdx.x];
    out[i] = v;
    if(i == n - 1) *total = v + in[i];
}

//! exclusive prefix sum of n values; *total receives the sum. block_sums holds div_up(n, 1024) entries
static void exclusive_scan(hipStream_t s, const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *block_sums,
                           uint32_t *total)
{
    const uint32_t nb = div_up(n, 1024);
    k_scan_local<<<nb, 1024, 0, s>>>(in, n, out, block_sums);
    k_scan_blocks<<<1, 1024, 0, s>>>(block_sums, nb);
    k_scan_add<<<nb, 1024, 0, s>>>(in, n, out, block_sums, total);
}

static void build_topology_ploc(vkrt_ctx *ctx, vkrt_bvh *bvh)
{
    hipStream_t s = ctx->stream;
    const uint32_t radius = bvh->ploc_radius ? bvh->ploc_radius : PLOC_RADIUS;
    const uint32_t n = bvh->n_tri, ni = n - 1;
    // working set: two cluster lists (ids, lo, hi), neighbours, flags and offsets, leaf counts, scan sums
    device_buffer &w = ctx->scratch[3];
    const size_t nb = div_up(n, 1024);
    const size_t bytes = (size_t)n * (2 * 4 + 4 * 16 + 4 + 4 * 4 + 4) + (size_t)ni * 4 + nb * 4 + 512;
    w.reserve(bytes);
    uint8_t *q = w.as<uint8_t>();
    auto take = [&q](size_t b) {
        uint8_t *r = q;
        q += (b + 15) & ~size_t(15);
        return r;
    };
    int32_t *id[2] = {reinterpret_cast<int32_t *>(take((size_t)n * 4)), reinterpret_cast<int32_t *>(take((size_t)n * 4))};
    float4 *lo[2] = {reinterpret_cast<float4 *>(take((size_t)n * 16)), reinterpret_cast<float4 *>(take((size_t)n * 16))};
    float4 *hi[2] = {reinterpret_cast<float4 *>(take((size_t)n * 16)), reinterpret_cast<float4 *>(take((size_t)n * 16))};
    int32_t *nn = reinterpret_cast<int32_t *>(take((size_t)n * 4));
    uint32_t *merge = reinterpret_cast<uint32_t *>(take((size_t)n * 4));
    uint32_t *moff = reinterpret_cast<uint32_t *>(take((size_t)n * 4));
    uint32_t *keep = reinterpret_cast<uint32_t *>(take((size_t)n * 4));
    uint32_t *koff = reinterpret_cast<uint32_t *>(take((size_t)n * 4));
    uint32_t *leaf_pos = reinterpret_cast<uint32_t *>(take((size_t)n * 4));
    uint32_t *cnt = reinterpret_cast<uint32_t *>(take((size_t)ni * 4));
    uint32_t *sums = reinterpret_cast<uint32_t *>(take(nb * 4));
    uint32_t *totals = reinterpret_cast<uint32_t *>(take(16));
    VKRT_REQUIRE(q <= w.as<uint8_t>() + bytes, "PLOC working set overrun");

    int2 *child = bvh->child.as<int2>();
    int32_t *parent_int = bvh->parent.as<int32_t>(), *parent_leaf = parent_int + ni;
    k_ploc_init<<<div_up(n, 256), 256, 0, s>>>(bvh->tbox.as<float>(), n, id[0], lo[0], hi[0], parent_int, ni);
    uint32_t nc = n, base = 0;
    int cur = 0;
    uint32_t rounds = 0;
    while(nc > 1)
    {
        VKRT_REQUIRE(++rounds <= n, "PLOC made no progress");
        k_ploc_nn<<<div_up(nc, PLOC_BLOCK), PLOC_BLOCK, 0, s>>>(lo[cur], hi[cur], nc, static_cast<int>(radius), nn);
        k_ploc_flags<<<div_up(nc, 256), 256, 0, s>>>(nn, nc, merge, keep);
        exclusive_scan(s, merge, moff, nc, sums, totals);
        exclusive_scan(s, keep, koff, nc, sums, totals + 1);
        k_ploc_merge<<<div_up(nc, 256), 256, 0, s>>>(nn, nc, merge, moff, keep, koff, base, id[cur], lo[cur], hi[cur],
                                                     id[cur ^ 1], lo[cur ^ 1], hi[cur ^ 1], child, cnt, parent_int,
                                                     parent_leaf);
        uint32_t t[2];
        VKRT_HIP(hipMemcpyAsync(t, totals, 8, hipMemcpyDeviceToHost, s));
        VKRT_HIP(hipStreamSynchronize(s));
        VKRT_REQUIRE(t[0] >= 1 && t[1] == nc - t[0], "PLOC round inconsistent");
        base += t[0];
        nc = t[1];
        cur ^= 1;
    }
    VKRT_REQUIRE(base == ni, "PLOC did not build n - 1 nodes");
    // the cluster lists and flag arrays are free now: renumbered tree (root first), then depth-first leaf
    // positions and ranges, then the permuted copies, which are copied back into the BVH's own arrays.
    // lo[1] and hi[0] are adjacent (32 n bytes), enough for the 24 n bytes of permuted boxes
    int2 *child2 = reinterpret_cast<int2 *>(lo[0]);
    uint32_t *cnt2 = merge;
    int32_t *pint2 = reinterpret_cast<int32_t *>(moff);
    int32_t *pleaf2 = reinterpret_cast<int32_t *>(keep);
    k_ploc_renumber<<<div_up(n, 256), 256, 0, s>>>(ni, n, child, cnt, parent_int, parent_leaf, child2, cnt2, pint2,
                                                   pleaf2);
    k_ploc_dfs<<<div_up(n, 256), 256, 0, s>>>(ni, n, child2, cnt2, pint2, pleaf2, leaf_pos, bvh->range.as<uint2>());
    uint32_t *order_out = koff;
    float *tbox_out = reinterpret_cast<float *>(lo[1]);
    int32_t *pleaf_out = id[0];
    int2 *child_out = reinterpret_cast<int2 *>(hi[1]);
    k_ploc_permute<<<div_up(n, 256), 256, 0, s>>>(ni, n, leaf_pos, bvh->order.as<uint32_t>(), bvh->tbox.as<float>(),
                                                  pleaf2, order_out, tbox_out, pleaf_out, child2, child_out);
    VKRT_HIP(hipMemcpyAsync(bvh->order.ptr, order_out, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    VKRT_HIP(hipMemcpyAsync(bvh->tbox.ptr, tbox_out, (size_t)n * 24, hipMemcpyDeviceToDevice, s));
    VKRT_HIP(hipMemcpyAsync(parent_leaf, pleaf_out, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    VKRT_HIP(hipMemcpyAsync(parent_int, pint2, (size_t)ni * 4, hipMemcpyDeviceToDevice, s));
    VKRT_HIP(hipMemcpyAsync(child, child_out, (size_t)ni * 8, hipMemcpyDeviceToDevice, s));
    VKRT_HIP(hipGetLastError());
}

static void build_bounds(vkrt_ctx *ctx, vkrt_bvh *bvh)
{
    hipStream_t s = ctx->stream;
    const uint32_t n = bvh->n_tri;
    if(n >= 2)
    {
        const uint32_t ni = n - 1;
        // scratch: per-leaf prefix and suffix unions, then the segments' sparse table
        const uint32_t nseg = div_up(n, BOX_SEG), levels = box_levels(nseg);
        device_buffer &w = ctx->scratch[3];
        w.reserve(((size_t)2 * n + (size_t)levels * nseg) * 6 * sizeof(float));
        float *pre = w.as<float>(), *suf = pre + 6 * (size_t)n, *table = suf + 6 * (size_t)n;
        k_box_scan<<<nseg, BOX_SEG, 0, s>>>(n, bvh->tbox.as<float>(), pre, suf, table);
        for(uint32_t k = 1; k < levels; ++k)
            k_box_sparse<<<div_up(nseg, 256), 256, 0, s>>>(nseg, k, table);
        k_box_nodes<<<div_up(ni, 256), 256, 0, s>>>(ni, n, nseg, bvh->range.as<uint2>(), bvh->tbox.as<float>(), pre,
                                                     suf, table, bvh->nbox.as<float>());
        bvh->n_nodes = ni;
        bvh->nodes.reserve((size_t)ni * sizeof(vkrt_bvh_node));
        k_layout<<<div_up(ni, 256), 256, 0, s>>>(n, bvh->child.as<int2>(), bvh->range.as<uint2>(),
                                                 bvh->nbox.as<float>(), bvh->tbox.as<float>(), bvh->max_leaf,
                                                 bvh->nodes.as<vkrt_bvh_node>());
        bvh->root_ref = (n <= bvh->max_leaf) ? ~static_cast<int>(n - 1) : 0;
    }
    else
    {
        bvh->n_nodes = 0;
        bvh->root_ref = ~0;// a leaf holding the single triangle
    }
    bvh->tris.reserve((size_t)n * 3 * sizeof(float4));
    k_reorder<<<div_up(n, 256), 256, 0, s>>>(bvh->order.as<uint32_t>(), n, bvh->pre.as<float>(),
                                             bvh->tris.as<float4>());
    VKRT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// 4-wide collapse: top-down, one BVH4 node per work item. A work item starts from a BVH2 node's two
// children and repeatedly opens the internal child with the largest surface area until it has four
// children or only leaves; each internal child becomes a new work item. The boxes are the BVH2's (already
// padded), so traversal stays conservative. Every item also carries the traversal stack depth on entry
// (the sum of pushes, children - 1, over its ancestors); the maximum decides whether the collapse fits the
// fixed stack.
// ---------------------------------------------------------------------------------------------
struct collapse_child
{
    int32_t ref;
    int32_t src;// (BVH2 parent node << 1) | side: where the box is stored, re-read by a refit
    float lo[3], hi[3];
};

__device__ __forceinline__ collapse_child child_of(const vkrt_bvh_node &nd, int32_t node, int k)
{
    collapse_child c;
    const float4 xy = k == 0 ? nd.a : nd.b;
    c.ref = k == 0 ? nd.d.x : nd.d.y;
    c.src = (node << 1) | k;
    c.lo[0] = xy.x;
    c.hi[0] = xy.y;
    c.lo[1] = xy.z;
    c.hi[1] = xy.w;
    c.lo[2] = k == 0 ? nd.c.x : nd.c.z;
    c.hi[2] = k == 0 ? nd.c.y : nd.c.w;
    return c;
}

__device__ __forceinline__ float half_area(const collapse_child &c)
{
    const float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

//! queue entries: (BVH2 node, BVH4 slot, stack depth on entry, -); counters: [0] BVH4 nodes allocated,
//! [1] next queue size, [2] maximum stack depth
__global__ void __launch_bounds__(256) k_collapse(const vkrt_bvh_node *__restrict__ n2, const int4 *__restrict__ in_q,
                                                  uint32_t n_in, vkrt_bvh4_node *__restrict__ n4,
                                                  int4 *__restrict__ out_q, uint32_t *__restrict__ counters)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n_in) return;
    const int4 item = in_q[t];
    collapse_child ch[4];
    const vkrt_bvh_node root = n2[item.x];
    ch[0] = child_of(root, item.x, 0);
    ch[1] = child_of(root, item.x, 1);
    int cnt = 2;
    while(cnt < 4)
    {
        int best = -1;
        float best_a = -1.0f;
        for(int j = 0; j < cnt; ++j)
        {
            if(ch[j].ref < 0) continue;
            const float a = half_area(ch[j]);
            if(a > best_a)
            {
                best_a = a;
                best = j;
            }
        }
        if(best < 0) break;
        const int32_t opened = ch[best].ref;
        const vkrt_bvh_node nd = n2[opened];
        ch[best] = child_of(nd, opened, 0);
        ch[cnt] = child_of(nd, opened, 1);
        ++cnt;
    }
    const int32_t depth_here = item.z + (cnt - 1);
    atomicMax(&counters[2], static_cast<uint32_t>(depth_here));
    vkrt_bvh4_node out;
    float lo[3][4], hi[3][4];
    int32_t refs[4], srcs[4];
    for(int j = 0; j < 4; ++j)
    {
        srcs[j] = j < cnt ? ch[j].src : -1;
        for(int a = 0; a < 3; ++a)
        {
            lo[a][j] = j < cnt ? ch[j].lo[a] : 0.0f;
            hi[a][j] = j < cnt ? ch[j].hi[a] : 0.0f;
        }
        refs[j] = EMPTY_CHILD;
        if(j >= cnt) continue;
        if(ch[j].ref < 0) refs[j] = ch[j].ref;
        else
        {
            const uint32_t slot = atomicAdd(&counters[0], 1u);
            refs[j] = static_cast<int32_t>(slot);
            const uint32_t q = atomicAdd(&counters[1], 1u);
            out_q[q] = make_int4(ch[j].ref, static_cast<int32_t>(slot), depth_here, 0);
        }
    }
    out.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    out.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    out.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    out.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    out.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    out.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    out.child = make_int4(refs[0], refs[1], refs[2], refs[3]);
    out.src = make_int4(srcs[0], srcs[1], srcs[2], srcs[3]);
    n4[item.y] = out;
}

//! refit of the 4-wide layout: every used child slot re-reads its box from the BVH2 node it was collapsed
//! from (the BVH2 keeps its topology and node numbering on refit), so the collapse's shape, its worst-case
//! stack depth and the choice of the 4-wide walk stay what the build recorded
__global__ void k_wide_refit(const vkrt_bvh_node *__restrict__ n2, uint32_t n4_count, vkrt_bvh4_node *__restrict__ n4)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n4_count) return;
    vkrt_bvh4_node nd = n4[i];
    float *lox = &nd.lox.x, *hix = &nd.hix.x, *loy = &nd.loy.x, *hiy = &nd.hiy.x, *loz = &nd.loz.x, *hiz = &nd.hiz.x;
    const int32_t src[4] = {nd.src.x, nd.src.y, nd.src.z, nd.src.w};
    for(int j = 0; j < 4; ++j)
    {
        if(src[j] < 0) continue;
        const vkrt_bvh_node p = n2[src[j] >> 1];
        const collapse_child c = child_of(p, src[j] >> 1, src[j] & 1);
        lox[j] = c.lo[0];
        hix[j] = c.hi[0];
        loy[j] = c.lo[1];
        hiy[j] = c.hi[1];
        loz[j] = c.lo[2];
        hiz[j] = c.hi[2];
    }
    n4[i] = nd;
}

//! refreshes the boxes of an existing 4-wide layout after the BVH2 was refitted (no change of shape)
static void refit_wide(vkrt_ctx *ctx, vkrt_bvh *bvh)
{
    if(bvh->empty || !bvh->wide || bvh->root4 < 0 || bvh->n_nodes4 == 0) return;
    k_wide_refit<<<div_up(bvh->n_nodes4, 256), 256, 0, ctx->stream>>>(bvh->nodes.as<vkrt_bvh_node>(), bvh->n_nodes4,
                                                                       bvh->nodes4.as<vkrt_bvh4_node>());
    VKRT_HIP(hipGetLastError());
}

//! collapse the BVH2 into the BVH4 used for traversal; keeps the BVH2 path when the worst-case stack
//! depth of the collapse exceeds STACK_SIZE
static void build_wide(vkrt_ctx *ctx, vkrt_bvh *bvh)
{
    hipStream_t s = ctx->stream;
    bvh->wide = false;
    bvh->n_nodes4 = 0;
    bvh->stack4 = 0;
    if(bvh->empty || !bvh->allow_wide) return;
    if(bvh->root_ref < 0)
    {
        bvh->root4 = bvh->root_ref;// a single leaf
        bvh->wide = true;
        return;
    }
    const size_t cap = std::max<uint32_t>(1u, bvh->n_nodes);
    bvh->nodes4.reserve(cap * sizeof(vkrt_bvh4_node));
    bvh->collapse_q[0].reserve(cap * sizeof(int4));
    bvh->collapse_q[1].reserve(cap * sizeof(int4));
    bvh->collapse_cnt.reserve(64);
    uint32_t *cnt = bvh->collapse_cnt.as<uint32_t>();
    const uint32_t init[4] = {1u, 0u, 0u, 0u};
    VKRT_HIP(hipMemcpyAsync(cnt, init, sizeof init, hipMemcpyHostToDevice, s));
    const int4 root_item = make_int4(bvh->root_ref, 0, 0, 0);
    VKRT_HIP(hipMemcpyAsync(bvh->collapse_q[0].ptr, &root_item, sizeof root_item, hipMemcpyHostToDevice, s));
    uint32_t n_in = 1;
    int cur = 0;
    while(n_in)
    {
        VKRT_HIP(hipMemsetAsync(cnt + 1, 0, 4, s));
        k_collapse<<<div_up(n_in, 256), 256, 0, s>>>(bvh->nodes.as<vkrt_bvh_node>(), bvh->collapse_q[cur].as<int4>(),
                                                     n_in, bvh->nodes4.as<vkrt_bvh4_node>(),
                                                     bvh->collapse_q[cur ^ 1].as<int4>(), cnt);
        VKRT_HIP(hipGetLastError());
        uint32_t next = 0;
        VKRT_HIP(hipMemcpyAsync(&next, cnt + 1, 4, hipMemcpyDeviceToHost, s));
        VKRT_HIP(hipStreamSynchronize(s));
        n_in = next;
        cur ^= 1;
    }
    uint32_t host_cnt[4];
    VKRT_HIP(hipMemcpyAsync(host_cnt, cnt, sizeof host_cnt, hipMemcpyDeviceToHost, s));
    VKRT_HIP(hipStreamSynchronize(s));
    bvh->n_nodes4 = host_cnt[0];
    bvh->stack4 = host_cnt[2];
    bvh->root4 = 0;
    VKRT_REQUIRE(bvh->n_nodes4 <= cap, "BVH4 collapse overflow");
    // the 4-wide walk addresses nodes with 32-bit byte offsets (node4_intervals)
    bvh->wide = bvh->stack4 <= (uint32_t)STACK_SIZE &&
                (uint64_t)bvh->n_nodes4 * sizeof(vkrt_bvh4_node) <= 0xffffffffull;
}

//! the info block after a build (stats: depth, leaf count and SAH cost from all nodes, read back to the host)
//! or after a refit, which keeps the topology: only the bounds change, read from the root node
static void refresh_info(vkrt_bvh *bvh, hipStream_t s, bool stats)
{
    bvh->info.n_tri = bvh->n_tri;
    bvh->info.n_nodes = bvh->n_nodes;
    bvh->info.device_bytes = bvh->nodes.bytes + bvh->tris.bytes + bvh->nodes4.bytes;
    bvh->info.n_wide_nodes = bvh->wide ? bvh->n_nodes4 : 0;
    bvh->info.wide_stack = bvh->stack4;
    if(stats || bvh->empty || bvh->root_ref < 0)
    {
        compute_stats(bvh, s);// (a single leaf or an empty BVH reads no node array)
        return;
    }
    vkrt_bvh_node root;
    VKRT_HIP(hipMemcpyAsync(&root, bvh->nodes.ptr, sizeof root, hipMemcpyDeviceToHost, s));
    VKRT_HIP(hipStreamSynchronize(s));
    float rb[6];
    root_bounds(root, rb, bvh->info.bounds);
}

}// namespace vkrt

using namespace vkrt;

extern "C" {

int vkrt_bvh_build(vkrt_ctx *ctx, const float *positions, uint32_t n_vert, uint32_t stride_bytes,
                   const uint32_t *indices, uint32_t n_tri, const vkrt_bvh_params *params, int flags, vkrt_bvh **out)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(out, "null output handle");
        *out = nullptr;
        VKRT_REQUIRE(stride_bytes >= 12 && stride_bytes % 4 == 0, "stride must be >= 12 and a multiple of 4");
        VKRT_REQUIRE(n_tri == 0 || (positions && indices && n_vert), "null mesh arrays");
        VKRT_REQUIRE(n_tri < (1u << 28), "at most 2^28 - 1 triangles per BVH");
        uint32_t leaf = params && params->max_leaf_size ? params->max_leaf_size : 2;
        VKRT_REQUIRE(leaf >= 1 && leaf <= (uint32_t)MAX_LEAF, "max_leaf_size must be in 1..8");
        VKRT_REQUIRE(!params || params->builder == VKRT_BUILDER_LBVH || params->builder == VKRT_BUILDER_PLOC,
                     "builder must be VKRT_BUILDER_LBVH or VKRT_BUILDER_PLOC");
        VKRT_REQUIRE(!params || params->width == 0 || params->width == 2 || params->width == 4,
                     "width must be 0 (auto), 2 or 4");
        auto t_start = std::chrono::steady_clock::now();
        auto bvh = std::make_unique<vkrt_bvh>();
        bvh->device = ctx->device;
        bvh->n_tri = n_tri;
        bvh->n_vert = n_vert;
        bvh->max_leaf = leaf;
        bvh->empty = n_tri == 0;
        bvh->allow_wide = !params || params->width != 2;
        bvh->builder = params ? params->builder : VKRT_BUILDER_LBVH;
        VKRT_REQUIRE(!params || params->ploc_radius <= (uint32_t)PLOC_MAX_RADIUS, "ploc_radius must be in 0..32");
        bvh->ploc_radius = params ? params->ploc_radius : 0u;
        if(n_tri)
        {
            const size_t pos_bytes = (size_t)(n_vert - 1) * stride_bytes + 12;
            const uint8_t *dpos =
                    input(ctx, 0, reinterpret_cast<const uint8_t *>(positions), pos_bytes, flags);
            bvh->indices.reserve((size_t)n_tri * 3 * sizeof(uint32_t));
            const uint32_t *didx = input(ctx, 1, indices, (size_t)n_tri * 3, flags);
            VKRT_HIP(hipMemcpyAsync(bvh->indices.ptr, didx, (size_t)n_tri * 3 * sizeof(uint32_t),
                                    hipMemcpyDeviceToDevice, ctx->stream));
            build_topology(ctx, bvh.get(), dpos, stride_bytes, bvh->indices.as<uint32_t>());
            build_bounds(ctx, bvh.get());
            build_wide(ctx, bvh.get());
            VKRT_HIP(hipStreamSynchronize(ctx->stream));
        }
        bvh->info.build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        refresh_info(bvh.get(), ctx->stream, true);
        *out = bvh.release();
    });
}

int vkrt_bvh_refit(vkrt_ctx *ctx, vkrt_bvh *bvh, const float *positions, uint32_t stride_bytes, int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(bvh, "null bvh");
        VKRT_REQUIRE(stride_bytes >= 12 && stride_bytes % 4 == 0, "stride must be >= 12 and a multiple of 4");
        if(bvh->empty) return;
        VKRT_REQUIRE(positions, "null positions");
        hipStream_t s = ctx->stream;
        const uint32_t n = bvh->n_tri;
        const size_t pos_bytes = (size_t)(bvh->n_vert - 1) * stride_bytes + 12;
        const uint8_t *dpos = input(ctx, 0, reinterpret_cast<const uint8_t *>(positions), pos_bytes, flags);
        device_buffer &tmp = ctx->scratch[2];
        tmp.reserve((size_t)n * 6 * sizeof(float) + 64);
        float *box = tmp.as<float>();
        uint32_t *cb = reinterpret_cast<uint32_t *>(tmp.as<uint8_t>() + (size_t)n * 6 * sizeof(float));
        VKRT_HIP(hipMemsetAsync(bvh->err.ptr, 0, 4, s));
        k_gather<<<div_up(n, 256), 256, 0, s>>>(dpos, stride_bytes, bvh->indices.as<uint32_t>(), n, bvh->n_vert,
                                                bvh->pre.as<float>(), box, cb, bvh->err.as<uint32_t>());
        k_sorted_boxes<<<div_up(n, 256), 256, 0, s>>>(bvh->order.as<uint32_t>(), n, box, bvh->tbox.as<float>());
        build_bounds(ctx, bvh);
        refit_wide(ctx, bvh);
        VKRT_HIP(hipStreamSynchronize(s));
        refresh_info(bvh, s, false);
    });
}

int vkrt_bvh_validate(vkrt_ctx *ctx, const vkrt_bvh *bvh, const float *positions, uint32_t stride_bytes, int flags,
                      vkrt_bvh_check *out)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(bvh && out, "null bvh or output");
        VKRT_REQUIRE(stride_bytes >= 12 && stride_bytes % 4 == 0, "stride must be >= 12 and a multiple of 4");
        VKRT_REQUIRE(bvh->empty || positions, "null positions");
        *out = vkrt_bvh_check{};
        out->wide = bvh->wide ? 1u : 0u;
        if(bvh->empty)
        {
            out->ok = 1;
            out->stack_ok = 1;
            return;
        }
        hipStream_t s = ctx->stream;
        const uint32_t n = bvh->n_tri;
        auto fetch = [&](const void *src, size_t bytes, bool device, std::vector<uint8_t> &dst) {
            dst.resize(bytes);
            if(!bytes) return;
            VKRT_HIP(hipMemcpyAsync(dst.data(), src, bytes, device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost, s));
        };
        std::vector<uint8_t> h_nodes, h_nodes4, h_tris, h_idx, h_pos;
        fetch(bvh->nodes.ptr, (size_t)bvh->n_nodes * sizeof(vkrt_bvh_node), true, h_nodes);
        if(bvh->wide) fetch(bvh->nodes4.ptr, (size_t)bvh->n_nodes4 * sizeof(vkrt_bvh4_node), true, h_nodes4);
        fetch(bvh->tris.ptr, (size_t)n * 3 * sizeof(float4), true, h_tris);
        fetch(bvh->indices.ptr, (size_t)n * 3 * sizeof(uint32_t), true, h_idx);
        fetch(positions, (size_t)(bvh->n_vert - 1) * stride_bytes + 12, (flags & VKRT_DEVICE_PTRS) != 0, h_pos);
        VKRT_HIP(hipStreamSynchronize(s));
        const auto *nodes = reinterpret_cast<const vkrt_bvh_node *>(h_nodes.data());
        const auto *nodes4 = reinterpret_cast<const vkrt_bvh4_node *>(h_nodes4.data());
        const auto *tris = reinterpret_cast<const float4 *>(h_tris.data());
        const auto *idx = reinterpret_cast<const uint32_t *>(h_idx.data());
        std::vector<uint32_t> seen(n, 0);
        auto vertex = [&](uint32_t v, int a) {
            float x;
            std::memcpy(&x, h_pos.data() + (size_t)v * stride_bytes + 4 * a, 4);
            return x;
        };
        auto violation = [&](uint32_t kind, uint32_t node, uint32_t slot, uint32_t tri, const float *box,
                             const float *outer, int n_outer) {
            if(out->box_violations++ != 0) return;
            out->bad_kind = kind;
            out->bad_node = node;
            out->bad_slot = slot;
            out->bad_tri = tri;
            std::memcpy(out->bad_box, box, sizeof out->bad_box);
            for(int a = 0; a < n_outer; ++a) out->bad_outer[a] = outer[a];
        };
        // a leaf's triangles: counted, and their vertices inside the box the parent stores for the leaf
        auto leaf = [&](int32_t ref, const float *box, uint32_t node, uint32_t slot) {
            const uint32_t v = static_cast<uint32_t>(~ref), first = v >> 3, count = (v & 7u) + 1;
            if((uint64_t)first + count > n)
            {
                ++out->bad_refs;
                return;
            }
            for(uint32_t k = first; k < first + count; ++k)
            {
                ++out->tris_reached;
                uint32_t orig;
                std::memcpy(&orig, &tris[3 * (size_t)k + 2].y, 4);
                if(orig >= n)
                {
                    ++out->bad_refs;
                    continue;
                }
                ++seen[orig];
                if(!box) continue;
                bool finite = true, inside = true;
                float outside[3] = {0.f, 0.f, 0.f};
                for(int c = 0; c < 3; ++c)
                {
                    const uint32_t vi = idx[3 * (size_t)orig + c];
                    if(vi >= bvh->n_vert)
                    {
                        finite = false;
                        break;
                    }
                    bool in = true;
                    for(int a = 0; a < 3; ++a)
                    {
                        const float x = vertex(vi, a);
                        finite = finite && std::isfinite(x);
                        in = in && x >= box[a] && x <= box[3 + a];
                    }
                    if(!in && inside)
                        for(int a = 0; a < 3; ++a) outside[a] = vertex(vi, a);
                    inside = inside && in;
                }
                // non-finite triangles get empty boxes
                if(finite && !inside) violation(2, node, slot, orig, box, outside, 3);
            }
        };
        auto inside = [](const float *in, const float *outer) {
            for(int a = 0; a < 3; ++a)
                if(!(in[a] >= outer[a] && in[3 + a] <= outer[3 + a])) return false;
            return true;
        };
        uint32_t recorded = 0;
        if(bvh->root_ref < 0)
        {
            leaf(bvh->root_ref, nullptr, 0, 0);// a single leaf: no box is stored for it
            out->stack_bound = 0;
        }
        else if(bvh->wide)
        {
            recorded = bvh->stack4;
            struct item
            {
                int32_t node;
                uint32_t depth;
                float box[6];
                bool has_box;
            };
            std::vector<item> todo{{0, 0u, {}, false}};
            while(!todo.empty())
            {
                const item it = todo.back();
                todo.pop_back();
                if(it.node < 0 || (uint32_t)it.node >= bvh->n_nodes4)
                {
                    ++out->bad_refs;
                    continue;
                }
                const vkrt_bvh4_node &nd = nodes4[it.node];
                const int32_t ch[4] = {nd.child.x, nd.child.y, nd.child.z, nd.child.w};
                const float *lx = &nd.lox.x, *hx = &nd.hix.x, *ly = &nd.loy.x, *hy = &nd.hiy.x, *lz = &nd.loz.x,
                            *hz = &nd.hiz.x;
                uint32_t used = 0;
                for(int j = 0; j < 4; ++j) used += ch[j] != EMPTY_CHILD ? 1u : 0u;
                const uint32_t depth_here = it.depth + (used ? used - 1 : 0);
                out->stack_bound = std::max(out->stack_bound, depth_here);
                for(int j = 0; j < 4; ++j)
                {
                    if(ch[j] == EMPTY_CHILD) continue;
                    const float b[6] = {lx[j], ly[j], lz[j], hx[j], hy[j], hz[j]};
                    if(it.has_box && !inside(b, it.box)) violation(1, it.node, j, 0, b, it.box, 6);
                    if(ch[j] < 0) leaf(ch[j], b, it.node, j);
                    else
                    {
                        item c{ch[j], depth_here, {}, true};
                        std::memcpy(c.box, b, sizeof b);
                        todo.push_back(c);
                    }
                }
            }
        }
        else
        {
            // binary tree: the stack holds at most one entry per ancestor of the deepest node
            recorded = bvh->info.depth;
            struct item
            {
                int32_t node;
                uint32_t depth;
                float box[6];
                bool has_box;
            };
            std::vector<item> todo{{0, 1u, {}, false}};
            while(!todo.empty())
            {
                const item it = todo.back();
                todo.pop_back();
                if(it.node < 0 || (uint32_t)it.node >= bvh->n_nodes)
                {
                    ++out->bad_refs;
                    continue;
                }
                const vkrt_bvh_node &nd = nodes[it.node];
                out->stack_bound = std::max(out->stack_bound, it.depth + 1);
                const float b[2][6] = {{nd.a.x, nd.a.z, nd.c.x, nd.a.y, nd.a.w, nd.c.y},
                                       {nd.b.x, nd.b.z, nd.c.z, nd.b.y, nd.b.w, nd.c.w}};
                const int32_t ch[2] = {nd.d.x, nd.d.y};
                for(int j = 0; j < 2; ++j)
                {
                    if(it.has_box && !inside(b[j], it.box)) violation(1, it.node, j, 0, b[j], it.box, 6);
                    if(ch[j] < 0) leaf(ch[j], b[j], it.node, j);
                    else
                    {
                        item c{ch[j], it.depth + 1, {}, true};
                        std::memcpy(c.box, b[j], sizeof b[j]);
                        todo.push_back(c);
                    }
                }
            }
        }
        for(uint32_t k = 0; k < n; ++k) out->tris_not_once += seen[k] != 1 ? 1u : 0u;
        out->stack_ok = (bvh->root_ref < 0 || out->stack_bound == recorded) && out->stack_bound <= (uint32_t)STACK_SIZE;
        out->ok = out->tris_not_once == 0 && out->box_violations == 0 && out->bad_refs == 0 && out->stack_ok &&
                  out->tris_reached == n;
    });
}

int vkrt_bvh_get_info(const vkrt_bvh *bvh, vkrt_bvh_info *out)
{
    if(!bvh || !out) return VKRT_E_INVALID_ARG;
    *out = bvh->info;
    return VKRT_OK;
}

void vkrt_bvh_destroy(vkrt_bvh *bvh)
{
    if(!bvh) return;
    (void)hipSetDevice(bvh->device);
    delete bvh;
}

}// extern "C"
