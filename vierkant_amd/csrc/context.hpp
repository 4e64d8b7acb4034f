// context.hpp - the vkrt_ctx behind the C ABI: device, stream, error string and staging buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "vkrt.h"

namespace vkrt
{

//! error carrying a C-ABI status code; thrown inside the library, caught at the boundary
struct error : std::runtime_error
{
    int code;
    error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

#define VKRT_HIP(expr)                                                                                         \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if(e_ != hipSuccess)                                                                                   \
            throw ::vkrt::error(e_ == hipErrorOutOfMemory ? VKRT_E_OOM : VKRT_E_HIP,                           \
                                std::string(#expr) + ": " + hipGetErrorString(e_));                           \
    } while(0)

#define VKRT_REQUIRE(cond, msg)                                                                                \
    do {                                                                                                       \
        if(!(cond)) throw ::vkrt::error(VKRT_E_INVALID_ARG, msg);                                             \
    } while(0)

//! grow-only device allocation
struct device_buffer
{
    void *ptr = nullptr;
    size_t bytes = 0;

    device_buffer() = default;
    device_buffer(const device_buffer &) = delete;
    device_buffer &operator=(const device_buffer &) = delete;
    device_buffer(device_buffer &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr, o.bytes = 0; }
    device_buffer &operator=(device_buffer &&o) noexcept
    {
        if(this != &o)
        {
            release();
            ptr = o.ptr;
            bytes = o.bytes;
            o.ptr = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~device_buffer() { release(); }

    void reserve(size_t n)
    {
        if(n <= bytes) return;
        release();
        size_t alloc = n < 256 ? 256 : n;
        VKRT_HIP(hipMalloc(&ptr, alloc));
        bytes = alloc;
    }
    void release()
    {
        if(ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template<class T>
    T *as() const
    {
        return static_cast<T *>(ptr);
    }
};

}// namespace vkrt

struct vkrt_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string last_error;
    int cu_count = 256;
    // fastest shape of the block x LDS depth x loop x XCD chunk sweep on MI355X (1.8M-triangle scene)
    vkrt_trace_config trace_cfg{64, 16, VKRT_TRAVERSAL_STEP, 32};
    //! staging slots for host-pointer calls (inputs/outputs copied through these)
    vkrt::device_buffer staging[8];
    //! scratch for multi-pass kernels (keys, counters)
    vkrt::device_buffer scratch[4];
    //! this device's address of the allocation-free entry point's file-scope scratch array
    void *noalloc_scratch = nullptr;
};

namespace vkrt
{

//! run `fn` with the context's device current; convert exceptions to status codes
template<class F>
int guarded(vkrt_ctx *ctx, F &&fn)
{
    if(!ctx) return VKRT_E_INVALID_ARG;
    try
    {
        VKRT_HIP(hipSetDevice(ctx->device));
        fn();
        ctx->last_error.clear();
        return VKRT_OK;
    } catch(const error &e)
    {
        ctx->last_error = e.what();
        return e.code;
    } catch(const std::bad_alloc &)
    {
        ctx->last_error = "host allocation failed";
        return VKRT_E_OOM;
    } catch(const std::exception &e)
    {
        ctx->last_error = e.what();
        return VKRT_E_INTERNAL;
    }
}

//! An input array: the device pointer to use, staged through ctx->staging[slot] for host pointers.
template<class T>
const T *input(vkrt_ctx *ctx, int slot, const T *p, size_t count, int flags)
{
    if(flags & VKRT_DEVICE_PTRS || count == 0) return p;
    auto &buf = ctx->staging[slot];
    buf.reserve(count * sizeof(T));
    VKRT_HIP(hipMemcpyAsync(buf.ptr, p, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return buf.as<const T>();
}

//! An output array: the device pointer kernels write to.
template<class T>
T *output(vkrt_ctx *ctx, int slot, T *p, size_t count, int flags)
{
    if(flags & VKRT_DEVICE_PTRS || count == 0) return p;
    auto &buf = ctx->staging[slot];
    buf.reserve(count * sizeof(T));
    return buf.as<T>();
}

//! Copy a staged output back to the host pointer (no-op for device pointers).
template<class T>
void finish_output(vkrt_ctx *ctx, int slot, T *host, size_t count, int flags)
{
    if(flags & VKRT_DEVICE_PTRS || count == 0) return;
    VKRT_HIP(hipMemcpyAsync(host, ctx->staging[slot].ptr, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
}

//! Synchronize unless the call is asynchronous; host-pointer calls always synchronize.
inline void finish_call(vkrt_ctx *ctx, int flags)
{
    if(!(flags & VKRT_ASYNC) || !(flags & VKRT_DEVICE_PTRS)) VKRT_HIP(hipStreamSynchronize(ctx->stream));
    VKRT_HIP(hipGetLastError());
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

}// namespace vkrt
