// trace.hip - closest/any-hit launches through a BVH, with a tunable launch shape.
//
// Replaces the RayTracer::trace_rays launch (src/RayTracer.cpp:66-135) and the driver's TraceRay /
// ray-query traversal (shaders/slang/raypipeline.slang:182, ray/visibility_test.slang:45-87). The kernels
// are instantiated for block sizes {64, 128, 256} and LDS stack depths {0, 8, 16, 32}; the context's
// trace configuration (vkrt_set_trace_config) selects one, with the loop of the 4-wide walk and the
// order in which workgroups take rays.
#include <chrono>
#include <vector>

#include "bvh_types.hpp"
#include "context.hpp"
#include "trace.hpp"
#include "traverse.hpp"

namespace vkrt
{

template<bool ANY, int BLOCK, int LDS_DEPTH, bool POSTPONE>
__global__ void __launch_bounds__(BLOCK) k_trace(const vkrt_bvh_node *__restrict__ nodes,
                                                 const float4 *__restrict__ tris, int32_t root, bool empty,
                                                 const vkrt_bvh4_node *__restrict__ nodes4, int32_t root4, bool wide,
                                                 const float4 *__restrict__ rays, uint32_t n, void *__restrict__ out,
                                                 uint32_t chunk)
{
    __shared__ int32_t smem[LDS_DEPTH > 0 ? LDS_DEPTH * BLOCK : 1];
    const uint32_t i = xcd_block(blockIdx.x, chunk) * BLOCK + threadIdx.x;
    if(i >= n) return;
    float t = 0.f, u = 0.f, v = 0.f;
    uint32_t prim = VKRT_INVALID_PRIM;
    bool hit = false;
    if(!empty)
    {
        ray_setup r = setup_ray(rays, i);
        if(!wide) hit = traverse<ANY, BLOCK, LDS_DEPTH>(nodes, tris, root, r, smem + threadIdx.x, t, u, v, prim);
        else if(POSTPONE)
            hit = traverse4_postponed<ANY, BLOCK, LDS_DEPTH>(nodes4, tris, root4, r, smem + threadIdx.x, t, u, v, prim);
        else
            hit = traverse4<ANY, BLOCK, LDS_DEPTH>(nodes4, tris, root4, r, smem + threadIdx.x, t, u, v, prim);
    }
    if(ANY) static_cast<uint8_t *>(out)[i] = hit ? 1 : 0;
    else
        static_cast<vkrt_hit *>(out)[i] = hit ? vkrt_hit{t, u, v, prim}
                                              : vkrt_hit{__int_as_float(0x7f800000), 0.f, 0.f, VKRT_INVALID_PRIM};
}

template<bool ANY, int BLOCK, int LDS, bool POSTPONE>
static void launch_one(hipStream_t s, const vkrt_bvh *bvh, const float4 *rays, uint32_t n, void *out, uint32_t chunk)
{
    k_trace<ANY, BLOCK, LDS, POSTPONE><<<xcd_grid(div_up(n, BLOCK), chunk), BLOCK, 0, s>>>(
            bvh->nodes.as<vkrt_bvh_node>(), bvh->tris.as<float4>(), bvh->root_ref, bvh->empty,
            bvh->nodes4.as<vkrt_bvh4_node>(), bvh->root4, bvh->wide, rays, n, out, chunk);
}

template<bool ANY, int BLOCK>
static void launch_block(vkrt_ctx *ctx, const vkrt_trace_config &cfg, const vkrt_bvh *bvh, const float4 *rays,
                         uint32_t n, void *out)
{
    const bool p = cfg.traversal == VKRT_TRAVERSAL_POSTPONED;
    const uint32_t x = cfg.xcd_chunk;
    hipStream_t s = ctx->stream;
    switch(cfg.lds_stack_depth)
    {
        case 0: p ? launch_one<ANY, BLOCK, 0, true>(s, bvh, rays, n, out, x)
                  : launch_one<ANY, BLOCK, 0, false>(s, bvh, rays, n, out, x); break;
        case 8: p ? launch_one<ANY, BLOCK, 8, true>(s, bvh, rays, n, out, x)
                  : launch_one<ANY, BLOCK, 8, false>(s, bvh, rays, n, out, x); break;
        case 16: p ? launch_one<ANY, BLOCK, 16, true>(s, bvh, rays, n, out, x)
                   : launch_one<ANY, BLOCK, 16, false>(s, bvh, rays, n, out, x); break;
        case 32: p ? launch_one<ANY, BLOCK, 32, true>(s, bvh, rays, n, out, x)
                   : launch_one<ANY, BLOCK, 32, false>(s, bvh, rays, n, out, x); break;
        default: throw error(VKRT_E_INVALID_ARG, "lds_stack_depth must be 0, 8, 16 or 32");
    }
}

bool valid_trace_config(const vkrt_trace_config &c)
{
    return (c.block_size == 64 || c.block_size == 128 || c.block_size == 256) &&
           (c.lds_stack_depth == 0 || c.lds_stack_depth == 8 || c.lds_stack_depth == 16 ||
            c.lds_stack_depth == 32) &&
           c.traversal <= VKRT_TRAVERSAL_POSTPONED && c.xcd_chunk <= 256 && (c.xcd_chunk & (c.xcd_chunk - 1)) == 0;
}

void launch_trace(vkrt_ctx *ctx, const vkrt_trace_config &cfg, bool any, const vkrt_bvh *bvh, const float4 *rays,
                  uint32_t n, void *out)
{
    if(!n) return;
    VKRT_REQUIRE(valid_trace_config(cfg), "unsupported trace configuration");
    switch(cfg.block_size)
    {
        case 64: any ? launch_block<true, 64>(ctx, cfg, bvh, rays, n, out)
                     : launch_block<false, 64>(ctx, cfg, bvh, rays, n, out); break;
        case 128: any ? launch_block<true, 128>(ctx, cfg, bvh, rays, n, out)
                      : launch_block<false, 128>(ctx, cfg, bvh, rays, n, out); break;
        default: any ? launch_block<true, 256>(ctx, cfg, bvh, rays, n, out)
                     : launch_block<false, 256>(ctx, cfg, bvh, rays, n, out); break;
    }
    VKRT_HIP(hipGetLastError());
}

}// namespace vkrt

using namespace vkrt;

extern "C" {

int vkrt_trace_closest(vkrt_ctx *ctx, const vkrt_bvh *bvh, const vkrt_ray *rays, uint32_t n, vkrt_hit *out,
                       int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(bvh, "null bvh");
        VKRT_REQUIRE(bvh->device == ctx->device, "bvh belongs to another device");
        VKRT_REQUIRE(n == 0 || (rays && out), "null array argument");
        if(!n) return;
        const vkrt_ray *dr = input(ctx, 1, rays, n, flags);
        vkrt_hit *dh = output(ctx, 2, out, n, flags);
        launch_trace(ctx, ctx->trace_cfg, false, bvh, reinterpret_cast<const float4 *>(dr), n, dh);
        finish_output(ctx, 2, out, n, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_trace_any(vkrt_ctx *ctx, const vkrt_bvh *bvh, const vkrt_ray *rays, uint32_t n, uint8_t *occluded,
                   int flags)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(bvh, "null bvh");
        VKRT_REQUIRE(bvh->device == ctx->device, "bvh belongs to another device");
        VKRT_REQUIRE(n == 0 || (rays && occluded), "null array argument");
        if(!n) return;
        const vkrt_ray *dr = input(ctx, 1, rays, n, flags);
        uint8_t *docc = output(ctx, 2, occluded, n, flags);
        launch_trace(ctx, ctx->trace_cfg, true, bvh, reinterpret_cast<const float4 *>(dr), n, docc);
        finish_output(ctx, 2, occluded, n, flags);
        finish_call(ctx, flags);
    });
}

int vkrt_set_trace_config(vkrt_ctx *ctx, const vkrt_trace_config *cfg)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(cfg && valid_trace_config(*cfg), "block_size must be 64/128/256 and lds_stack_depth 0/8/16/32");
        ctx->trace_cfg = *cfg;
    });
}

int vkrt_get_trace_config(const vkrt_ctx *ctx, vkrt_trace_config *cfg)
{
    if(!ctx || !cfg) return VKRT_E_INVALID_ARG;
    *cfg = ctx->trace_cfg;
    return VKRT_OK;
}

}// extern "C"
