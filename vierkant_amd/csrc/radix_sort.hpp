// radix_sort.hpp - project-written stable LSD radix sort of (uint32 key, uint32 value) pairs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace vkrt
{

//! device scratch bytes radix_sort_pairs needs for n pairs
size_t radix_sort_temp_bytes(uint32_t n);

//! sort keys (low key_bits significant) with their values in place, stable; temp from radix_sort_temp_bytes
void radix_sort_pairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t key_bits, void *temp);

}// namespace vkrt
