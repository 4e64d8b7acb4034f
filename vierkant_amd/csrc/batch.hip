// batch.hip - many scenes and ray batches of uneven size traced by one persistent launch, replayed as a
// captured hipGraph, with a launch-shape autotuner.
//
// The reference traces one scene per vkCmdTraceRaysKHR (src/RayTracer.cpp:128-129); this batches the work
// of many independent scenes. Every scene's rays are cut into tiles of `tile` rays; a device table holds
// each item's BVH and buffers and the prefix of its tile counts. Blocks of k_batch_trace claim tiles with
// one atomic each (from per-XCD queues, see claim_tile) and leave once every queue is exhausted, so every
// wave reaches the exit. The graph is [k_batch_reset (zeroes the counters with a kernel), k_batch_trace];
// sizes and pointers live in device memory, so the captured graph replays without re-capture until the
// launch shape changes.
#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <memory>
#include <vector>

#include "bvh_types.hpp"
#include "context.hpp"
#include "traverse.hpp"

namespace vkrt
{

struct batch_item_dev
{
    const vkrt_bvh_node *nodes;
    const float4 *tris;
    const float4 *rays;
    vkrt_hit *hits;
    const vkrt_bvh4_node *nodes4;
    int32_t root;
    uint32_t empty;
    uint32_t n_rays;
    int32_t root4;
    uint32_t wide;
    uint32_t pad[3];
};

constexpr uint32_t N_COUNTERS = CLAIM_COUNTERS;

__global__ void k_batch_reset(uint32_t *counter) { counter[threadIdx.x] = 0u; }

//! (8 waves per SIMD where the LDS stack allows it: the traversal fits 64 VGPRs with a spill or two of the
//! loop state, against 72 unconstrained, which holds 7)
template<int BLOCK, int LDS>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8))) k_batch_trace(const batch_item_dev *__restrict__ items, uint32_t n_items,
                                                       const uint32_t *__restrict__ tile_prefix, uint32_t tile,
                                                       uint32_t *__restrict__ counters, uint32_t chunk)
{
    __shared__ int32_t smem[LDS > 0 ? LDS * BLOCK : 1];
    __shared__ uint32_t s_tile;
    const uint32_t total = tile_prefix[n_items];
    const uint32_t home = blockIdx.x % XCDS;
    uint32_t queue = 0;// thread 0's position in the queue order (only thread 0 claims)
    while(true)
    {
        if(threadIdx.x == 0) s_tile = claim_tile(counters, chunk, home, total, queue);
        __syncthreads();
        // every lane reads the same word: make it scalar, so the item lookup and its fields stay in SGPRs
        const uint32_t t = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if(t >= total) break;
        // item owning tile t: the last item whose prefix is <= t (binary search, block-uniform)
        uint32_t lo = 0, hi = n_items;
        while(hi - lo > 1)
        {
            const uint32_t mid = (lo + hi) >> 1;
            if(tile_prefix[mid] <= t) lo = mid;
            else hi = mid;
        }
        const batch_item_dev it = items[lo];
        const uint32_t first = (t - tile_prefix[lo]) * tile;
        const uint32_t count = min(tile, it.n_rays - first);
        for(uint32_t k = threadIdx.x; k < count; k += BLOCK)
        {
            const uint32_t r = first + k;
            float bt = 0.f, bu = 0.f, bv = 0.f;
            uint32_t prim = VKRT_INVALID_PRIM;
            bool hit = false;
            if(!it.empty)
            {
                ray_setup rs = setup_ray(it.rays, r);
                hit = it.wide ? traverse4<false, BLOCK, LDS>(it.nodes4, it.tris, it.root4, rs, smem + threadIdx.x, bt,
                                                             bu, bv, prim)
                              : traverse<false, BLOCK, LDS>(it.nodes, it.tris, it.root, rs, smem + threadIdx.x, bt, bu,
                                                            bv, prim);
            }
            it.hits[r] = hit ? vkrt_hit{bt, bu, bv, prim}
                             : vkrt_hit{__int_as_float(0x7f800000), 0.f, 0.f, VKRT_INVALID_PRIM};
        }
    }
}

template<int BLOCK, int LDS>
static void launch_batch(hipStream_t s, uint32_t grid, const batch_item_dev *items, uint32_t n_items,
                         const uint32_t *prefix, uint32_t tile, uint32_t *counter, uint32_t chunk)
{
    k_batch_trace<BLOCK, LDS><<<grid, BLOCK, 0, s>>>(items, n_items, prefix, tile, counter, chunk);
}

template<int BLOCK>
static void launch_batch_lds(hipStream_t s, uint32_t lds, uint32_t grid, const batch_item_dev *items,
                             uint32_t n_items, const uint32_t *prefix, uint32_t tile, uint32_t *counter,
                             uint32_t chunk)
{
    switch(lds)
    {
        case 0: launch_batch<BLOCK, 0>(s, grid, items, n_items, prefix, tile, counter, chunk); break;
        case 8: launch_batch<BLOCK, 8>(s, grid, items, n_items, prefix, tile, counter, chunk); break;
        case 16: launch_batch<BLOCK, 16>(s, grid, items, n_items, prefix, tile, counter, chunk); break;
        case 24: launch_batch<BLOCK, 24>(s, grid, items, n_items, prefix, tile, counter, chunk); break;
        case 32: launch_batch<BLOCK, 32>(s, grid, items, n_items, prefix, tile, counter, chunk); break;
        case 48: launch_batch<BLOCK, 48>(s, grid, items, n_items, prefix, tile, counter, chunk); break;
        default: launch_batch<BLOCK, 64>(s, grid, items, n_items, prefix, tile, counter, chunk); break;
    }
}

constexpr size_t LDS_PER_CU = 160 * 1024;

static bool valid_batch_config(const vkrt_batch_config &c)
{
    const bool block = c.block_size == 64 || c.block_size == 128 || c.block_size == 256;
    const bool lds = c.lds_stack_depth == 0 || c.lds_stack_depth == 8 || c.lds_stack_depth == 16 ||
                     c.lds_stack_depth == 24 || c.lds_stack_depth == 32 || c.lds_stack_depth == 48 ||
                     c.lds_stack_depth == 64;
    const bool tile = c.tile >= 64 && c.tile <= (1u << 20) && (c.tile % 64) == 0;
    const bool chunk = c.xcd_chunk <= 64;
    // one block's LDS stack must fit the CU's 160 KiB
    const bool fits = (size_t)c.lds_stack_depth * c.block_size * 4 <= LDS_PER_CU;
    return block && lds && tile && chunk && fits && c.blocks_per_cu >= 1 && c.blocks_per_cu <= 64;
}

// the autotuner's shape space: block sizes, LDS-resident stack depths over the 16-64 range (the traversal
// stack is 64 entries; the rest lives in scratch), tiles, resident threads per CU (blocks per CU scale with
// the block size) and the XCD chunk of the claim order. Shapes whose blocks cannot all hold their LDS stack
// at once on a CU are left out.
constexpr uint32_t TUNE_BLOCKS[] = {64, 128, 256};
constexpr uint32_t TUNE_LDS[] = {16, 32, 48, 64};
constexpr uint32_t TUNE_TILES[] = {256, 1024, 4096};
constexpr uint32_t TUNE_THREADS_PER_CU[] = {512, 1024, 2048};
constexpr uint32_t TUNE_CHUNKS[] = {0, 4};

constexpr vkrt_batch_config DEFAULT_CONFIG{256, 16, 1024, 8, 4, {0, 0, 0}};// one of the candidates

static std::vector<vkrt_batch_config> candidate_shapes()
{
    std::vector<vkrt_batch_config> out;
    for(uint32_t bs: TUNE_BLOCKS)
        for(uint32_t ld: TUNE_LDS)
            for(uint32_t tl: TUNE_TILES)
                for(uint32_t th: TUNE_THREADS_PER_CU)
                    for(uint32_t ch: TUNE_CHUNKS)
                    {
                        const uint32_t per_cu = th / bs;
                        if((size_t)per_cu * ld * bs * 4 > LDS_PER_CU) continue;
                        vkrt_batch_config c{bs, ld, tl, per_cu, ch, {0, 0, 0}};
                        if(valid_batch_config(c)) out.push_back(c);
                    }
    return out;
}

}// namespace vkrt

struct vkrt_batch
{
    vkrt_ctx *ctx = nullptr;
    std::vector<vkrt::batch_item_dev> items;
    std::vector<uint32_t> n_rays;
    vkrt::device_buffer d_items, d_prefix, d_counter;
    vkrt_batch_config cfg = vkrt::DEFAULT_CONFIG;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t total_rays = 0;
};

namespace vkrt
{

static void release_graph(vkrt_batch *b)
{
    if(b->exec) (void)hipGraphExecDestroy(b->exec);
    if(b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

//! tile prefix for the current tile size, uploaded, then the graph captured
static void prepare(vkrt_batch *b)
{
    hipStream_t s = b->ctx->stream;
    const uint32_t n = static_cast<uint32_t>(b->items.size());
    std::vector<uint32_t> prefix(n + 1, 0);
    for(uint32_t i = 0; i < n; ++i)
    {
        const uint64_t tiles = (b->n_rays[i] + (uint64_t)b->cfg.tile - 1) / b->cfg.tile;
        VKRT_REQUIRE(prefix[i] + tiles < 0xffffffffull, "too many tiles");
        prefix[i + 1] = prefix[i] + static_cast<uint32_t>(tiles);
    }
    b->d_prefix.reserve(prefix.size() * 4);
    VKRT_HIP(hipMemcpyAsync(b->d_prefix.ptr, prefix.data(), prefix.size() * 4, hipMemcpyHostToDevice, s));
    VKRT_HIP(hipStreamSynchronize(s));
    release_graph(b);
    const uint32_t grid = std::max<uint32_t>(1, b->cfg.blocks_per_cu * b->ctx->cu_count);
    VKRT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    k_batch_reset<<<1, N_COUNTERS, 0, s>>>(b->d_counter.as<uint32_t>());
    const batch_item_dev *items = b->d_items.as<batch_item_dev>();
    const uint32_t *pre = b->d_prefix.as<uint32_t>();
    uint32_t *counter = b->d_counter.as<uint32_t>();
    const uint32_t ld = b->cfg.lds_stack_depth, tl = b->cfg.tile, ch = b->cfg.xcd_chunk;
    switch(b->cfg.block_size)
    {
        case 64: launch_batch_lds<64>(s, ld, grid, items, n, pre, tl, counter, ch); break;
        case 128: launch_batch_lds<128>(s, ld, grid, items, n, pre, tl, counter, ch); break;
        default: launch_batch_lds<256>(s, ld, grid, items, n, pre, tl, counter, ch); break;
    }
    hipError_t launch_err = hipGetLastError();
    hipError_t end_err = hipStreamEndCapture(s, &b->graph);
    VKRT_HIP(launch_err);
    VKRT_HIP(end_err);
    VKRT_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

static float timed_launch(vkrt_batch *b, int repeats)
{
    hipStream_t s = b->ctx->stream;
    hipEvent_t e0, e1;
    VKRT_HIP(hipEventCreate(&e0));
    VKRT_HIP(hipEventCreate(&e1));
    VKRT_HIP(hipGraphLaunch(b->exec, s));// warm-up
    VKRT_HIP(hipEventRecord(e0, s));
    for(int k = 0; k < repeats; ++k) VKRT_HIP(hipGraphLaunch(b->exec, s));
    VKRT_HIP(hipEventRecord(e1, s));
    VKRT_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    VKRT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / repeats;
}

}// namespace vkrt

using namespace vkrt;

extern "C" {

int vkrt_batch_create(vkrt_ctx *ctx, const vkrt_batch_item *items, uint32_t n_items, vkrt_batch **out)
{
    return guarded(ctx, [&] {
        VKRT_REQUIRE(out, "null output handle");
        *out = nullptr;
        VKRT_REQUIRE(n_items > 0 && items, "a batch needs at least one item");
        auto b = std::make_unique<vkrt_batch>();
        b->ctx = ctx;
        for(uint32_t i = 0; i < n_items; ++i)
        {
            const vkrt_batch_item &it = items[i];
            VKRT_REQUIRE(it.bvh, "null bvh in batch item");
            VKRT_REQUIRE(it.bvh->device == ctx->device, "bvh belongs to another device");
            VKRT_REQUIRE(it.n_rays == 0 || (it.rays && it.hits), "null ray or hit buffer in batch item");
            batch_item_dev d{};
            d.nodes = it.bvh->nodes.as<vkrt_bvh_node>();
            d.tris = it.bvh->tris.as<float4>();
            d.rays = reinterpret_cast<const float4 *>(it.rays);
            d.hits = it.hits;
            d.nodes4 = it.bvh->nodes4.as<vkrt_bvh4_node>();
            d.root = it.bvh->root_ref;
            d.empty = it.bvh->empty ? 1u : 0u;
            d.n_rays = it.n_rays;
            d.root4 = it.bvh->root4;
            d.wide = it.bvh->wide ? 1u : 0u;
            b->items.push_back(d);
            b->n_rays.push_back(it.n_rays);
            b->total_rays += it.n_rays;
        }
        b->d_items.reserve(b->items.size() * sizeof(batch_item_dev));
        VKRT_HIP(hipMemcpyAsync(b->d_items.ptr, b->items.data(), b->items.size() * sizeof(batch_item_dev),
                                hipMemcpyHostToDevice, ctx->stream));
        b->d_counter.reserve(N_COUNTERS * sizeof(uint32_t));
        prepare(b.get());
        *out = b.release();
    });
}

void vkrt_batch_destroy(vkrt_batch *b)
{
    if(!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    release_graph(b);
    delete b;
}

int vkrt_batch_set_config(vkrt_batch *b, const vkrt_batch_config *cfg)
{
    if(!b) return VKRT_E_INVALID_ARG;
    return guarded(b->ctx, [&] {
        VKRT_REQUIRE(cfg && valid_batch_config(*cfg),
                     "block_size 64/128/256, lds_stack_depth 0/8/16/24/32/48/64 (at most 160 KiB per block), tile a "
                     "multiple of 64, blocks_per_cu 1..64, xcd_chunk 0..64");
        b->cfg = *cfg;
        prepare(b);
    });
}

int vkrt_batch_get_config(const vkrt_batch *b, vkrt_batch_config *out)
{
    if(!b || !out) return VKRT_E_INVALID_ARG;
    *out = b->cfg;
    return VKRT_OK;
}

int vkrt_batch_launch(vkrt_batch *b, int flags)
{
    if(!b) return VKRT_E_INVALID_ARG;
    return guarded(b->ctx, [&] {
        VKRT_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
        if(!(flags & VKRT_ASYNC)) VKRT_HIP(hipStreamSynchronize(b->ctx->stream));
    });
}

int vkrt_batch_autotune(vkrt_batch *b, uint32_t repeats, vkrt_batch_config *best, float *best_ms)
{
    if(!b) return VKRT_E_INVALID_ARG;
    return guarded(b->ctx, [&] {
        // coordinate descent over the candidate shapes: starting from the current shape, each round moves to
        // the fastest candidate that differs from the winner in one field, until a round changes nothing
        const std::vector<vkrt_batch_config> cands = candidate_shapes();
        const int reps = repeats ? static_cast<int>(repeats) : 2;
        auto key = [](const vkrt_batch_config &c) {
            return std::array<uint32_t, 5>{c.block_size, c.lds_stack_depth, c.tile, c.blocks_per_cu, c.xcd_chunk};
        };
        std::map<std::array<uint32_t, 5>, float> timed;
        auto measure = [&](const vkrt_batch_config &c) {
            auto it = timed.find(key(c));
            if(it != timed.end()) return it->second;
            b->cfg = c;
            prepare(b);
            const float ms = timed_launch(b, reps);
            timed[key(c)] = ms;
            return ms;
        };
        // the descent starts from the current shape when it is a candidate, else from the default one
        vkrt_batch_config win = DEFAULT_CONFIG;
        for(const vkrt_batch_config &c: cands)
            if(key(c) == key(b->cfg)) win = b->cfg;
        float win_ms = measure(win);
        for(int round = 0; round < 4; ++round)
        {
            bool moved = false;
            for(const vkrt_batch_config &c: cands)
            {
                const auto a = key(c), w = key(win);
                int diff = 0;
                for(int f = 0; f < 5; ++f) diff += a[f] != w[f];
                // blocks per CU follows the block size: a change of block size keeps the threads per CU
                const bool same_threads = c.block_size * c.blocks_per_cu == win.block_size * win.blocks_per_cu;
                const bool one_step = diff == 1 || (diff == 2 && a[0] != w[0] && a[3] != w[3] && same_threads);
                if(!one_step) continue;
                const float ms = measure(c);
                if(ms < win_ms)
                {
                    win_ms = ms;
                    win = c;
                    moved = true;
                }
            }
            if(!moved) break;
        }
        b->cfg = win;
        prepare(b);
        if(best) *best = win;
        if(best_ms) *best_ms = win_ms;
    });
}

uint32_t vkrt_batch_candidates(vkrt_batch_config *out, uint32_t capacity)
{
    const std::vector<vkrt_batch_config> c = candidate_shapes();
    for(uint32_t i = 0; out && i < capacity && i < c.size(); ++i) out[i] = c[i];
    return static_cast<uint32_t>(c.size());
}

uint64_t vkrt_batch_total_rays(const vkrt_batch *b) { return b ? b->total_rays : 0u; }

int vkrt_batch_closest(vkrt_ctx *ctx, const vkrt_batch_item *items, uint32_t n_items, int flags)
{
    vkrt_batch *b = nullptr;
    int rc = vkrt_batch_create(ctx, items, n_items, &b);
    if(rc != VKRT_OK) return rc;
    rc = vkrt_batch_launch(b, flags & ~VKRT_ASYNC);
    vkrt_batch_destroy(b);
    return rc;
}

}// extern "C"
