// scene.hip - two-level scenes: a top-level 4-wide BVH over instances of per-mesh BVHs (BLAS), each instance
// with its own transform.
//
// Replaces the top-level acceleration structure of RayBuilder (create_toplevel, src/RayBuilder.cpp:386-781;
// the instances' entry_t transform / inv_transform, include/vierkant/RayBuilder.hpp:24-38) and the driver's
// instance traversal. Every instance carries both its object-to-world and world-to-object matrices, like
// entry_t; traversal uses the world-to-object one, so the object-space ray is a fixed fp32 expression of the
// caller's numbers and the result is defined exactly (oracle: orc_scene_closest).
//
// The top level is small (one leaf per instance) and rebuilt on the host whenever transforms change:
// instance world boxes in double precision, padded, a 4-wide median split, uploaded as vkrt_bvh4_node. One
// lane traces one ray: top-level nodes exactly like a BLAS, and at an instance leaf the ray is transformed
// and that BLAS is walked on the same stack, above the top-level entries.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "bvh_types.hpp"
#include "context.hpp"
#include "kernels.hpp"
#include "scene_walk.hpp"

namespace vkrt
{

constexpr int SCENE_BLOCK = 64;   // with the context's XCD chunk: 3.02 Grays/s vs 2.96 at 128 and 2.85 at 256
constexpr int SCENE_LDS = 16;

template<bool ANY, int BLOCK, int LDS_DEPTH>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5))) k_scene_trace(const vkrt_bvh4_node *__restrict__ tlas, int32_t troot,
                                                       uint32_t empty_scene, const uint32_t *__restrict__ order,
                                                       const inst_dev *__restrict__ insts,
                                                       const blas_dev *__restrict__ blases,
                                                       const float4 *__restrict__ rays, uint32_t n,
                                                       void *__restrict__ out, uint32_t chunk)
{
    __shared__ int32_t smem[LDS_DEPTH > 0 ? LDS_DEPTH * BLOCK : 1];
    const uint32_t i = xcd_block(blockIdx.x, chunk) * BLOCK + threadIdx.x;
    if(i >= n) return;
    const float4 ra = rays[2 * (size_t)i], rb = rays[2 * (size_t)i + 1];
    float best_t = rb.w, best_u = 0.f, best_v = 0.f;
    uint32_t best_prim = VKRT_INVALID_PRIM, best_inst = VKRT_INVALID_PRIM;
    if(!empty_scene)
        scene_walk<ANY, BLOCK, LDS_DEPTH>(tlas, troot, order, insts, blases, ra, rb, smem + threadIdx.x, best_t,
                                          best_u, best_v, best_prim, best_inst);
    const bool hit = best_prim != VKRT_INVALID_PRIM;
    if(ANY) static_cast<uint8_t *>(out)[i] = hit ? 1 : 0;
    else
    {
        vkrt_instance_hit h;
        h.t = hit ? best_t : __int_as_float(0x7f800000);
        h.u = hit ? best_u : 0.f;
        h.v = hit ? best_v : 0.f;
        h.prim = best_prim;
        h.instance = best_inst;
        h.custom_index = hit ? insts[best_inst].custom : 0u;
        h.reserved[0] = h.reserved[1] = 0u;
        static_cast<vkrt_instance_hit *>(out)[i] = h;
    }
}

}// namespace vkrt

//! the scene object behind the C ABI
struct vkrt_scene
{
    vkrt_ctx *ctx = nullptr;
    std::vector<const vkrt_bvh *> blas;
    std::vector<vkrt_instance> instances;
    vkrt::device_buffer nodes, order, insts, blases;
    int32_t root = 0;
    bool empty = true;
    uint32_t n_nodes = 0, tlas_stack = 0, max_blas_stack = 0;
};

namespace vkrt
{

struct host_box
{
    double lo[3], hi[3];
};

//! conservative float of a double bound (rounded down for lower, up for upper bounds)
static float round_down(double x)
{
    float f = static_cast<float>(x);
    if(static_cast<double>(f) > x) f = std::nextafter(f, -INFINITY);
    return f;
}
static float round_up(double x)
{
    float f = static_cast<float>(x);
    if(static_cast<double>(f) < x) f = std::nextafter(f, INFINITY);
    return f;
}

struct tlas_builder
{
    const std::vector<host_box> &box;
    std::vector<uint32_t> &order;
    std::vector<vkrt_bvh4_node> &nodes;
    bool sah = true;
    uint32_t max_stack = 0;

    static host_box merge(const host_box &a, const host_box &b)
    {
        host_box r;
        for(int k = 0; k < 3; ++k)
        {
            r.lo[k] = std::min(a.lo[k], b.lo[k]);
            r.hi[k] = std::max(a.hi[k], b.hi[k]);
        }
        return r;
    }

    host_box bounds(uint32_t lo, uint32_t hi) const
    {
        host_box r = box[order[lo]];
        for(uint32_t j = lo + 1; j < hi; ++j) r = merge(r, box[order[j]]);
        return r;
    }

    static double half_area(const host_box &b)
    {
        const double dx = std::max(0.0, b.hi[0] - b.lo[0]), dy = std::max(0.0, b.hi[1] - b.lo[1]),
                     dz = std::max(0.0, b.hi[2] - b.lo[2]);
        return dx * dy + dy * dz + dz * dx;
    }

    //! split [lo, hi) into two non-empty parts by the surface area heuristic: for every axis the instances are
    //! put in SAH_BINS bins by box centre and the cut between bins with the least
    //! area(left) * count(left) + area(right) * count(right) wins; the range is then partitioned at that cut.
    //! Instances whose centres coincide on every axis (or a cut that would leave a side empty) fall back to
    //! the median of the widest centre axis
    uint32_t split(uint32_t lo, uint32_t hi)
    {
        constexpr int SAH_BINS = 32;
        double cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
        for(uint32_t j = lo; j < hi; ++j)
            for(int k = 0; k < 3; ++k)
            {
                const double c = 0.5 * (box[order[j]].lo[k] + box[order[j]].hi[k]);
                cl[k] = std::min(cl[k], c);
                ch[k] = std::max(ch[k], c);
            }
        auto bin_of = [&](uint32_t id, int k) {
            const double c = 0.5 * (box[id].lo[k] + box[id].hi[k]);
            const int b = static_cast<int>((c - cl[k]) / (ch[k] - cl[k]) * SAH_BINS);
            return std::min(std::max(b, 0), SAH_BINS - 1);
        };
        double best_cost = INFINITY;
        int best_axis = -1, best_cut = 0;
        for(int k = 0; k < 3; ++k)
        {
            if(!(ch[k] > cl[k])) continue;
            host_box bb[SAH_BINS];
            uint32_t bc[SAH_BINS] = {};
            for(auto &b: bb)
                for(int a = 0; a < 3; ++a)
                {
                    b.lo[a] = INFINITY;
                    b.hi[a] = -INFINITY;
                }
            for(uint32_t j = lo; j < hi; ++j)
            {
                const int b = bin_of(order[j], k);
                bb[b] = bc[b] ? merge(bb[b], box[order[j]]) : box[order[j]];
                ++bc[b];
            }
            // suffix areas and counts, then a sweep from the left
            double right_area[SAH_BINS];
            uint32_t right_count[SAH_BINS];
            host_box acc = bb[0];
            uint32_t cnt = 0;
            for(int b = SAH_BINS - 1; b >= 0; --b)
            {
                if(bc[b])
                {
                    acc = cnt ? merge(acc, bb[b]) : bb[b];
                    cnt += bc[b];
                }
                right_area[b] = cnt ? half_area(acc) : 0.0;
                right_count[b] = cnt;
            }
            host_box left = bb[0];
            uint32_t lcnt = 0;
            for(int b = 0; b + 1 < SAH_BINS; ++b)
            {
                if(bc[b])
                {
                    left = lcnt ? merge(left, bb[b]) : bb[b];
                    lcnt += bc[b];
                }
                if(!lcnt || !right_count[b + 1]) continue;
                const double cost = half_area(left) * lcnt + right_area[b + 1] * right_count[b + 1];
                if(cost < best_cost)
                {
                    best_cost = cost;
                    best_axis = k;
                    best_cut = b + 1;
                }
            }
        }
        if(sah && best_axis >= 0)
        {
            const auto first_right = std::partition(order.begin() + lo, order.begin() + hi, [&](uint32_t id) {
                return bin_of(id, best_axis) < best_cut;
            });
            const uint32_t mid = static_cast<uint32_t>(first_right - order.begin());
            if(mid > lo && mid < hi) return mid;
        }
        int axis = 0;
        for(int k = 1; k < 3; ++k)
            if(ch[k] - cl[k] > ch[axis] - cl[axis]) axis = k;
        const uint32_t mid = lo + (hi - lo) / 2;
        std::nth_element(order.begin() + lo, order.begin() + mid, order.begin() + hi, [&](uint32_t a, uint32_t b) {
            const double ca = box[a].lo[axis] + box[a].hi[axis], cb = box[b].lo[axis] + box[b].hi[axis];
            return ca < cb || (ca == cb && a < b);
        });
        return mid;
    }

    //! ref of the subtree over order[lo, hi): a leaf of one instance or a new 4-wide node
    int32_t build(uint32_t lo, uint32_t hi, uint32_t stack_in)
    {
        if(hi - lo == 1) return static_cast<int32_t>(~(lo << 3));
        uint32_t ranges[4][2];
        int cnt = 0;
        const uint32_t mid = split(lo, hi);
        const uint32_t halves[2][2] = {{lo, mid}, {mid, hi}};
        for(const auto &h: halves)
        {
            if(h[1] - h[0] >= 2)
            {
                const uint32_t m = split(h[0], h[1]);
                ranges[cnt][0] = h[0];
                ranges[cnt++][1] = m;
                ranges[cnt][0] = m;
                ranges[cnt++][1] = h[1];
            }
            else
            {
                ranges[cnt][0] = h[0];
                ranges[cnt++][1] = h[1];
            }
        }
        const uint32_t here = stack_in + static_cast<uint32_t>(cnt - 1);
        max_stack = std::max(max_stack, here);
        const size_t slot = nodes.size();
        nodes.emplace_back();
        float lo_[3][4], hi_[3][4];
        int32_t ch[4] = {EMPTY_CHILD, EMPTY_CHILD, EMPTY_CHILD, EMPTY_CHILD};
        for(int j = 0; j < 4; ++j)
            for(int k = 0; k < 3; ++k) lo_[k][j] = hi_[k][j] = 0.f;
        for(int j = 0; j < cnt; ++j)
        {
            const host_box b = bounds(ranges[j][0], ranges[j][1]);
            for(int k = 0; k < 3; ++k)
            {
                lo_[k][j] = round_down(b.lo[k]);
                hi_[k][j] = round_up(b.hi[k]);
            }
            ch[j] = build(ranges[j][0], ranges[j][1], here);
        }
        vkrt_bvh4_node &nd = nodes[slot];
        nd.lox = make_float4(lo_[0][0], lo_[0][1], lo_[0][2], lo_[0][3]);
        nd.hix = make_float4(hi_[0][0], hi_[0][1], hi_[0][2], hi_[0][3]);
        nd.loy = make_float4(lo_[1][0], lo_[1][1], lo_[1][2], lo_[1][3]);
        nd.hiy = make_float4(hi_[1][0], hi_[1][1], hi_[1][2], hi_[1][3]);
        nd.loz = make_float4(lo_[2][0], lo_[2][1], lo_[2][2], lo_[2][3]);
        nd.hiz = make_float4(hi_[2][0], hi_[2][1], hi_[2][2], hi_[2][3]);
        nd.child = make_int4(ch[0], ch[1], ch[2], ch[3]);
        nd.src = make_int4(-1, -1, -1, -1);
        return static_cast<int32_t>(slot);
    }
};

//! world box of an instance: the BLAS's (padded) root bounds through object_to_world in double, grown by
//! 2^-14 of the box's magnitude so that the rounding of world_to_object (about 2^-24 relative) cannot move a
//! hit of the object-space walk outside it
static bool instance_box(const vkrt_instance &in, const vkrt_bvh *b, host_box &out)
{
    if(b->empty) return false;
    const float *bb = b->info.bounds;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double mag = 0.0;
    for(int c = 0; c < 8; ++c)
    {
        const double p[3] = {(c & 1) ? bb[3] : bb[0], (c & 2) ? bb[4] : bb[1], (c & 4) ? bb[5] : bb[2]};
        for(int a = 0; a < 3; ++a)
        {
            const float *m = in.object_to_world + 4 * a;
            const double w = (double)m[0] * p[0] + (double)m[1] * p[1] + (double)m[2] * p[2] + (double)m[3];
            lo[a] = std::min(lo[a], w);
            hi[a] = std::max(hi[a], w);
            mag = std::max(mag, std::fabs(w));
            mag = std::max(mag, std::fabs((double)m[3]));
        }
    }
    const double pad = std::ldexp(mag, -14) + 1e-30;
    for(int a = 0; a < 3; ++a)
    {
        if(!std::isfinite(lo[a]) || !std::isfinite(hi[a])) return false;
        out.lo[a] = lo[a] - pad;
        out.hi[a] = hi[a] + pad;
    }
    return true;
}

static void build_tlas(vkrt_scene *sc)
{
    const uint32_t n = static_cast<uint32_t>(sc->instances.size());
    std::vector<host_box> box(n);
    std::vector<uint32_t> order;
    for(uint32_t k = 0; k < n; ++k)
        if(instance_box(sc->instances[k], sc->blas[sc->instances[k].blas], box[k])) order.push_back(k);
    std::vector<vkrt_bvh4_node> nodes;
    sc->empty = order.empty();
    sc->tlas_stack = 0;
    if(!sc->empty)
    {
        tlas_builder tb{box, order, nodes};
        nodes.reserve(order.size());
        sc->root = tb.build(0, static_cast<uint32_t>(order.size()), 0);
        sc->tlas_stack = tb.max_stack;
        if(sc->tlas_stack + sc->max_blas_stack > (uint32_t)SCENE_STACK)
        {
            // an SAH tree too deep for the stack: the median split's balanced tree instead
            nodes.clear();
            tlas_builder mb{box, order, nodes};
            mb.sah = false;
            sc->root = mb.build(0, static_cast<uint32_t>(order.size()), 0);
            sc->tlas_stack = mb.max_stack;
        }
    }
    VKRT_REQUIRE(sc->tlas_stack + sc->max_blas_stack <= (uint32_t)SCENE_STACK,
                 "scene too deep for the traversal stack (top level + deepest mesh)");
    hipStream_t s = sc->ctx->stream;
    sc->n_nodes = static_cast<uint32_t>(nodes.size());
    sc->nodes.reserve(std::max<size_t>(1, nodes.size()) * sizeof(vkrt_bvh4_node));
    sc->order.reserve(std::max<size_t>(1, order.size()) * 4);
    std::vector<inst_dev> in(std::max<uint32_t>(1, n));
    for(uint32_t k = 0; k < n; ++k)
    {
        const float *w = sc->instances[k].world_to_object;
        in[k].w0 = make_float4(w[0], w[1], w[2], w[3]);
        in[k].w1 = make_float4(w[4], w[5], w[6], w[7]);
        in[k].w2 = make_float4(w[8], w[9], w[10], w[11]);
        in[k].blas = sc->instances[k].blas;
        in[k].custom = sc->instances[k].custom_index;
        in[k].pad0 = in[k].pad1 = 0;
    }
    sc->insts.reserve(in.size() * sizeof(inst_dev));
    if(!nodes.empty())
        VKRT_HIP(hipMemcpyAsync(sc->nodes.ptr, nodes.data(), nodes.size() * sizeof(vkrt_bvh4_node),
                                hipMemcpyHostToDevice, s));
    if(!order.empty())
        VKRT_HIP(hipMemcpyAsync(sc->order.ptr, order.data(), order.size() * 4, hipMemcpyHostToDevice, s));
    VKRT_HIP(hipMemcpyAsync(sc->insts.ptr, in.data(), in.size() * sizeof(inst_dev), hipMemcpyHostToDevice, s));
    VKRT_HIP(hipStreamSynchronize(s));
}

static void check_instances(const vkrt_scene *sc, const vkrt_instance *inst, uint32_t n)
{
    VKRT_REQUIRE(n == 0 || inst, "null instance array");
    for(uint32_t k = 0; k < n; ++k)
    {
        VKRT_REQUIRE(inst[k].blas < sc->blas.size(), "instance mesh index out of range");
        for(int j = 0; j < 12; ++j)
            VKRT_REQUIRE(std::isfinite(inst[k].object_to_world[j]) && std::isfinite(inst[k].world_to_object[j]),
                         "instance transforms must be finite");
    }
}

scene_view scene_device_view(const vkrt_scene *sc)
{
    return scene_view{sc->nodes.as<vkrt_bvh4_node>(), sc->root, sc->empty ? 1u : 0u, sc->order.as<uint32_t>(),
                      sc->insts.as<inst_dev>(), sc->blases.as<blas_dev>()};
}

template<bool ANY>
static void launch_scene(vkrt_scene *sc, const float4 *rays, uint32_t n, void *out)
{
    // the workgroup order of the context's trace configuration (XCD chunks)
    const uint32_t chunk = sc->ctx->trace_cfg.xcd_chunk;
    k_scene_trace<ANY, SCENE_BLOCK, SCENE_LDS><<<xcd_grid(div_up(n, SCENE_BLOCK), chunk), SCENE_BLOCK, 0,
                                                 sc->ctx->stream>>>(
            sc->nodes.as<vkrt_bvh4_node>(), sc->root, sc->empty ? 1u : 0u, sc->order.as<uint32_t>(),
            sc->insts.as<inst_dev>(), sc->blases.as<blas_dev>(), rays, n, out, chunk);
    VKRT_HIP(hipGetLastError());
}

}// namespace vkrt

using namespace vkrt;

extern "C" {

int vkrt_scene_create(vkrt_ctx *ctx, vkrt_bvh *const *meshes, uint32_t n_meshes, const vkrt_instance *instances,
                      uint32_t n_instances, vkrt_scene **out)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(out, "null output handle");
        *out = nullptr;
        VKRT_REQUIRE(n_meshes == 0 || meshes, "null mesh list");
        auto sc = std::make_unique<vkrt_scene>();
        sc->ctx = ctx;
        std::vector<blas_dev> bd(std::max<uint32_t>(1, n_meshes));
        for(uint32_t m = 0; m < n_meshes; ++m)
        {
            const vkrt_bvh *b = meshes[m];
            VKRT_REQUIRE(b, "null mesh BVH");
            VKRT_REQUIRE(b->device == ctx->device, "mesh BVH belongs to another device");
            VKRT_REQUIRE(b->empty || b->wide, "mesh BVHs of a scene need the 4-wide layout (width 0 or 4)");
            sc->blas.push_back(b);
            bd[m] = blas_dev{b->nodes4.as<vkrt_bvh4_node>(), b->tris.as<float4>(), b->root4, b->empty ? 1u : 0u};
            sc->max_blas_stack = std::max(sc->max_blas_stack, b->stack4);
        }
        sc->blases.reserve(bd.size() * sizeof(blas_dev));
        VKRT_HIP(hipMemcpyAsync(sc->blases.ptr, bd.data(), bd.size() * sizeof(blas_dev), hipMemcpyHostToDevice,
                                ctx->stream));
        check_instances(sc.get(), instances, n_instances);
        sc->instances.assign(instances, instances + n_instances);
        build_tlas(sc.get());
        *out = sc.release();
    });
}

int vkrt_scene_set_instances(vkrt_scene *sc, const vkrt_instance *instances, uint32_t n_instances)
{
    if(!sc) return VKRT_E_INVALID_ARG;
    return guarded(sc->ctx, [&] {
        check_instances(sc, instances, n_instances);
        sc->instances.assign(instances, instances + n_instances);
        build_tlas(sc);
    });
}

void vkrt_scene_destroy(vkrt_scene *sc)
{
    if(!sc) return;
    (void)hipSetDevice(sc->ctx->device);
    (void)hipStreamSynchronize(sc->ctx->stream);
    delete sc;
}

int vkrt_scene_trace_closest(vkrt_ctx *ctx, vkrt_scene *sc, const vkrt_ray *rays, uint32_t n, vkrt_instance_hit *out,
                             int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(sc && sc->ctx == ctx, "null scene or scene of another context");
        VKRT_REQUIRE(n == 0 || (rays && out), "null array argument");
        if(!n) return;
        const vkrt_ray *dr = input(ctx, 1, rays, n, flags);
        vkrt_instance_hit *dh = output(ctx, 2, out, n, flags);
        launch_scene<false>(sc, reinterpret_cast<const float4 *>(dr), n, dh);
        finish_output(ctx, 2, out, n, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_scene_trace_any(vkrt_ctx *ctx, vkrt_scene *sc, const vkrt_ray *rays, uint32_t n, uint8_t *occluded, int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(sc && sc->ctx == ctx, "null scene or scene of another context");
        VKRT_REQUIRE(n == 0 || (rays && occluded), "null array argument");
        if(!n) return;
        const vkrt_ray *dr = input(ctx, 1, rays, n, flags);
        uint8_t *docc = output(ctx, 2, occluded, n, flags);
        launch_scene<true>(sc, reinterpret_cast<const float4 *>(dr), n, docc);
        finish_output(ctx, 2, occluded, n, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_scene_get_info(const vkrt_scene *sc, uint32_t *n_instances, uint32_t *n_nodes, uint32_t *stack_bound)
{
    if(!sc) return VKRT_E_INVALID_ARG;
    if(n_instances) *n_instances = static_cast<uint32_t>(sc->instances.size());
    if(n_nodes) *n_nodes = sc->n_nodes;
    if(stack_bound) *stack_bound = sc->tlas_stack + sc->max_blas_stack;
    return VKRT_OK;
}

void vkrt_instance_set_transform(vkrt_instance *inst, const float object_to_world[12])
{
    if(!inst || !object_to_world) return;
    double m[3][4];
    for(int r = 0; r < 3; ++r)
        for(int c = 0; c < 4; ++c)
        {
            m[r][c] = object_to_world[4 * r + c];
            inst->object_to_world[4 * r + c] = object_to_world[4 * r + c];
        }
    // inverse of the affine map in double: A^-1 by cofactors, translation -A^-1 t, rounded once to float
    const double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0],
                 h = m[2][1], k = m[2][2];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    const double inv[3][3] = {{A / det, -(b * k - c * h) / det, (b * f - c * e) / det},
                              {B / det, (a * k - c * g) / det, -(a * f - c * d) / det},
                              {C / det, -(a * h - b * g) / det, (a * e - b * d) / det}};
    for(int r = 0; r < 3; ++r)
    {
        const double tr = -(inv[r][0] * m[0][3] + inv[r][1] * m[1][3] + inv[r][2] * m[2][3]);
        for(int cc = 0; cc < 3; ++cc) inst->world_to_object[4 * r + cc] = static_cast<float>(inv[r][cc]);
        inst->world_to_object[4 * r + 3] = static_cast<float>(tr);
    }
}

}// extern "C"
