// trace.hpp - launch of the BVH traversal kernels with a given launch shape.
#pragma once

#include <hip/hip_runtime.h>

#include "vkrt.h"

struct vkrt_ctx;

namespace vkrt
{

bool valid_trace_config(const vkrt_trace_config &c);

//! closest (any = false, out = vkrt_hit[n]) or any hit (any = true, out = uint8_t[n]) for n device rays
void launch_trace(vkrt_ctx *ctx, const vkrt_trace_config &cfg, bool any, const vkrt_bvh *bvh, const float4 *rays,
                  uint32_t n, void *out);

}// namespace vkrt
