// context.cpp - C-ABI context lifetime, errors and streams.
#include "context.hpp"
#include "kernels.hpp"

extern "C" {

int vkrt_version(void) { return VKRT_VERSION_MAJOR * 1000 + VKRT_VERSION_MINOR; }

int vkrt_device_count(void)
{
    int n = 0;
    if(hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int vkrt_create(int device, void *hip_stream, vkrt_ctx **out)
{
    if(!out) return VKRT_E_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if(hipGetDeviceCount(&n) != hipSuccess || n == 0) return VKRT_E_NO_DEVICE;
    if(device < 0 || device >= n) return VKRT_E_INVALID_ARG;
    auto *ctx = new(std::nothrow) vkrt_ctx;
    if(!ctx) return VKRT_E_OOM;
    ctx->device = device;
    int rc = vkrt::guarded(ctx, [&] {
        if(hip_stream) ctx->stream = static_cast<hipStream_t>(hip_stream);
        else
        {
            VKRT_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            ctx->own_stream = true;
        }
        hipDeviceProp_t prop{};
        VKRT_HIP(hipGetDeviceProperties(&prop, device));
        ctx->cu_count = prop.multiProcessorCount;
        std::string arch = prop.gcnArchName;
        if(arch.rfind("gfx950", 0) != 0)
            throw vkrt::error(VKRT_E_UNSUPPORTED, "libvkrt is built for gfx950 only; device is " + arch);
        vkrt::noalloc_scratch_init(ctx);
    });
    if(rc != VKRT_OK)
    {
        if(ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return VKRT_OK;
}

void vkrt_destroy(vkrt_ctx *ctx)
{
    if(!ctx) return;
    (void)hipSetDevice(ctx->device);
    if(ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for(auto &b: ctx->staging) b.release();
    for(auto &b: ctx->scratch) b.release();
    if(ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *vkrt_last_error(const vkrt_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int vkrt_sync(vkrt_ctx *ctx)
{
    return vkrt::guarded(ctx, [&] {
        VKRT_HIP(hipStreamSynchronize(ctx->stream));
        VKRT_HIP(hipGetLastError());
    });
}

void *vkrt_stream(const vkrt_ctx *ctx) { return ctx ? static_cast<void *>(ctx->stream) : nullptr; }

int vkrt_device_alloc(vkrt_ctx *ctx, size_t bytes, void **out)
{
    return vkrt::guarded(ctx, [&] {
        VKRT_REQUIRE(out, "null output pointer");
        *out = nullptr;
        VKRT_HIP(hipMalloc(out, bytes ? bytes : 1));
    });
}

int vkrt_device_free(vkrt_ctx *ctx, void *ptr)
{
    return vkrt::guarded(ctx, [&] {
        if(ptr) VKRT_HIP(hipFree(ptr));
    });
}

int vkrt_copy(vkrt_ctx *ctx, void *dst, const void *src, size_t bytes, int kind, int flags)
{
    return vkrt::guarded(ctx, [&] {
        VKRT_REQUIRE(kind >= 0 && kind <= 2, "copy kind must be 0 (h2d), 1 (d2h) or 2 (d2d)");
        if(!bytes) return;
        VKRT_REQUIRE(dst && src, "null copy pointer");
        const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                          : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
        VKRT_HIP(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
        if(!(flags & VKRT_ASYNC)) VKRT_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}// extern "C"
