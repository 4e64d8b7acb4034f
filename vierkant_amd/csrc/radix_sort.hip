// radix_sort.hip - stable LSD radix sort of (uint32 key, uint32 value) pairs, 8-bit digits.
//
// One pass per digit, three kernels per pass:
//   k_rs_hist    per-tile 256-bin histogram in LDS, stored digit-major (hist[d * tiles + tile])
//   k_rs_scan    exclusive scan of the digit-major histogram (one workgroup, 8 counters per thread, chunked
//                with a carry)
//   k_rs_scatter stable scatter: within a round of 256 items a wave64 finds the lanes holding the same
//                digit with 8 ballots (match-any), ranks by popcount below its lane, and waves are
//                ordered through per-wave digit counts in LDS
// Tiles are 256 threads x 8 rounds; item k of thread t is tile_base + 256 k + t, so loads are coalesced
// and the in-tile order (round, thread) is the input order, which keeps the sort stable.
#include "context.hpp"
#include "radix_sort.hpp"

namespace vkrt
{

constexpr uint32_t RS_THREADS = 256, RS_ROUNDS = 8, RS_TILE = RS_THREADS * RS_ROUNDS, RS_BINS = 256;

__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint32_t *__restrict__ keys, uint32_t n, uint32_t shift,
                                                        uint32_t tiles, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t h[RS_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
    for(uint32_t k = 0; k < RS_ROUNDS; ++k)
    {
        uint32_t i = base + k * RS_THREADS + threadIdx.x;
        if(i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

//! items per thread of k_rs_scan: a 1.8 M-key sort has 225,536 counters, 28 steps of 8,192 instead of 221 of
//! 1,024 (each step waits on its loads and four barriers: 211 us per pass at one item per thread)
constexpr uint32_t RS_SCAN_ITEMS = 8;

__global__ void __launch_bounds__(1024) k_rs_scan(uint32_t *__restrict__ data, uint32_t n)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if(threadIdx.x == 0) carry = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for(uint32_t base = 0; base < n; base += 1024 * RS_SCAN_ITEMS)
    {
        // each thread takes RS_SCAN_ITEMS consecutive counters; the block scans the threads' sums
        const uint32_t first = base + threadIdx.x * RS_SCAN_ITEMS;
        uint32_t v[RS_SCAN_ITEMS], x = 0;
#pragma unroll
        for(uint32_t k = 0; k < RS_SCAN_ITEMS; ++k)
        {
            v[k] = first + k < n ? data[first + k] : 0u;
            x += v[k];
        }
        // inclusive wave scan
        uint32_t s = x;
        for(int off = 1; off < 64; off <<= 1)
        {
            uint32_t y = __shfl_up(s, off);
            if(lane >= (uint32_t)off) s += y;
        }
        if(lane == 63) wsum[wave] = s;
        __syncthreads();
        if(threadIdx.x < 16)
        {
            uint32_t w = wsum[threadIdx.x], acc = w;
            for(int off = 1; off < 16; off <<= 1)
            {
                uint32_t y = __shfl_up(acc, off, 16);
                if(threadIdx.x >= (uint32_t)off) acc += y;
            }
            wsum[threadIdx.x] = acc - w;// exclusive over waves
        }
        __syncthreads();
        const uint32_t c = carry;
        uint32_t run = c + wsum[wave] + s - x;// exclusive prefix of this thread's first counter
#pragma unroll
        for(uint32_t k = 0; k < RS_SCAN_ITEMS; ++k)
        {
            if(first + k < n) data[first + k] = run;
            run += v[k];
        }
        __syncthreads();
        if(threadIdx.x == 1023) carry = c + wsum[wave] + s;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint32_t *__restrict__ keys_in,
                                                           const uint32_t *__restrict__ vals_in, uint32_t n,
                                                           uint32_t shift, uint32_t tiles,
                                                           const uint32_t *__restrict__ offsets,
                                                           uint32_t *__restrict__ keys_out,
                                                           uint32_t *__restrict__ vals_out)
{
    __shared__ uint32_t gbase[RS_BINS];
    __shared__ uint32_t wcount[RS_THREADS / 64][RS_BINS];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    gbase[t] = offsets[t * tiles + blockIdx.x];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * RS_TILE;
    for(uint32_t k = 0; k < RS_ROUNDS; ++k)
    {
        for(uint32_t w = 0; w < RS_THREADS / 64; ++w) wcount[w][t] = 0;
        __syncthreads();
        const uint32_t i = base + k * RS_THREADS + t;
        const bool valid = i < n;
        uint32_t key = 0, val = 0, d = 0;
        if(valid)
        {
            key = keys_in[i];
            val = vals_in[i];
            d = (key >> shift) & 0xffu;
        }
        uint64_t peers = __ballot(valid);
        for(int b = 0; b < 8; ++b)
        {
            const bool bit = (d >> b) & 1u;
            const uint64_t ones = __ballot(valid && bit);
            peers &= bit ? ones : ~ones;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        if(valid && rank == 0) wcount[wave][d] = __popcll(peers);
        __syncthreads();
        if(valid)
        {
            uint32_t off = gbase[d] + rank;
            for(uint32_t w = 0; w < wave; ++w) off += wcount[w][d];
            keys_out[off] = key;
            vals_out[off] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for(uint32_t w = 0; w < RS_THREADS / 64; ++w) add += wcount[w][t];
        gbase[t] += add;
        __syncthreads();
    }
}

size_t radix_sort_temp_bytes(uint32_t n)
{
    const size_t tiles = div_up(n, RS_TILE);
    return tiles * RS_BINS * sizeof(uint32_t) + 2 * (size_t)n * sizeof(uint32_t) + 256;
}

void radix_sort_pairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t key_bits, void *temp)
{
    if(n <= 1) return;
    const uint32_t tiles = div_up(n, RS_TILE);
    uint32_t *hist = static_cast<uint32_t *>(temp);
    uint32_t *k_alt = hist + (size_t)tiles * RS_BINS;
    uint32_t *v_alt = k_alt + n;
    uint32_t *kin = keys, *vin = vals, *kout = k_alt, *vout = v_alt;
    const uint32_t passes = div_up(key_bits, 8);
    for(uint32_t p = 0; p < passes; ++p)
    {
        const uint32_t shift = 8 * p;
        k_rs_hist<<<tiles, RS_THREADS, 0, s>>>(kin, n, shift, tiles, hist);
        k_rs_scan<<<1, 1024, 0, s>>>(hist, tiles * RS_BINS);
        k_rs_scatter<<<tiles, RS_THREADS, 0, s>>>(kin, vin, n, shift, tiles, hist, kout, vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if(kin != keys)
    {
        VKRT_HIP(hipMemcpyAsync(keys, kin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        VKRT_HIP(hipMemcpyAsync(vals, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
}

}// namespace vkrt
