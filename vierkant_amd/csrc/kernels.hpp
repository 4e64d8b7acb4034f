// kernels.hpp - device helpers and launchers shared between the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "vkrt.h"

struct vkrt_ctx;

namespace vkrt
{

//! IEEE float -> uint32 whose unsigned order is the float order (-0 sorts below +0; callers that
//! need -0 == +0 pass t + 0.0f, which maps -0 to +0)
__host__ __device__ __forceinline__ uint32_t ordered_bits(float f)
{
    uint32_t b;
    __builtin_memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

//! 64-bit closest-hit key: smaller t first, then the smaller primitive index
__host__ __device__ __forceinline__ uint64_t hit_key(float t, uint32_t prim)
{
    return (static_cast<uint64_t>(ordered_bits(t + 0.0f)) << 32) | prim;
}

//! XCD-chunked workgroup order: the dispatcher hands workgroups to the 8 XCDs round robin, so workgroup
//! b = (g C + j) 8 + x runs on XCD x; it takes work block g 8 C + x C + j, i.e. each XCD takes C consecutive
//! blocks per turn and neighbouring rays share its L2. chunk 0 keeps b. A bijection on [0, xcd_grid(n, C)).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t chunk)
{
    if(!chunk) return b;
    const uint32_t x = b % 8u, j = (b / 8u) % chunk, g = b / (8u * chunk);
    return g * 8u * chunk + x * chunk + j;
}

//! workgroups to launch for `blocks` work blocks: whole rounds of 8 C with a chunk (the surplus ones return)
inline uint32_t xcd_grid(uint32_t blocks, uint32_t chunk)
{
    return chunk ? static_cast<uint32_t>((blocks + 8ull * chunk - 1) / (8ull * chunk) * 8ull * chunk) : blocks;
}

constexpr uint32_t XCDS = 8;           // MI355X: workgroup b is dispatched to XCD b % 8
constexpr uint32_t COUNTER_STRIDE = 16;// one 64-B line per queue counter
constexpr uint32_t CLAIM_COUNTERS = XCDS * COUNTER_STRIDE;// words of one set of claim counters

//! Tile claims of a persistent kernel from `counters` (CLAIM_COUNTERS words, zeroed before the launch).
//! chunk == 0: one queue, tiles in order. chunk == C > 0: the tile sequence is cut into chunks of C
//! consecutive tiles dealt round robin to the 8 XCDs; XCD x's i-th claim is tile ((i / C) 8 + x) C + i % C,
//! which grows with i, so a queue whose claim passes the tile count is exhausted for good. A caller first
//! drains its own XCD's queue (`home`; neighbouring tiles then share that XCD's L2), then the others in turn,
//! so no XCD idles while tiles remain, and leaves after at most one failed claim per queue. `queue` is the
//! caller's position in that order (starts at 0). Returns `total` when every queue is exhausted.
//! A queue is read before it is claimed from: counters only grow, so a value whose tile is past the end means
//! the queue is exhausted for good, and the caller moves on without an atomic (thousands of workgroups
//! leaving a persistent launch would otherwise queue up failed atomics on eight addresses).
__device__ __forceinline__ uint64_t claim_slot_tile(uint64_t i, uint32_t x, uint32_t chunk)
{
    return chunk ? ((i / chunk) * XCDS + x) * chunk + i % chunk : i;
}

__device__ __forceinline__ uint32_t claim_tile(uint32_t *__restrict__ counters, uint32_t chunk, uint32_t home,
                                               uint32_t total, uint32_t &queue)
{
    const uint32_t n_queues = chunk ? XCDS : 1u;
    while(queue < n_queues)
    {
        const uint32_t x = (home + queue) % n_queues;
        uint32_t *c = &counters[x * COUNTER_STRIDE];
        const uint32_t seen = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if(claim_slot_tile(seen, x, chunk) < total)
        {
            const uint64_t t = claim_slot_tile(atomicAdd(c, 1u), x, chunk);
            if(t < total) return static_cast<uint32_t>(t);
        }
        ++queue;
    }
    return total;
}

void launch_tri_prep(hipStream_t s, const float *tris, uint32_t n, float *pre);
void launch_fill_u64(hipStream_t s, uint64_t *p, uint32_t n, uint64_t value);
void launch_fill_u8(hipStream_t s, uint8_t *p, uint32_t n, uint8_t value);

//! brute-force closest hit over pre-processed triangles (v0, e1, e2); keys: n_ray scratch words
void brute_closest(vkrt_ctx *ctx, const float *pre, uint32_t n_tri, const float4 *rays, uint32_t n_ray,
                   uint64_t *keys, vkrt_hit *hits);
//! resolves the allocation-free scratch array on the context's device (once per device)
void noalloc_scratch_init(vkrt_ctx *ctx);

void brute_any(vkrt_ctx *ctx, const float *pre, uint32_t n_tri, const float4 *rays, uint32_t n_ray, uint8_t *occ);

}// namespace vkrt
