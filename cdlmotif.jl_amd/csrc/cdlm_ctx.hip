// cdlm_ctx.hip -- context, stream and error reporting of libcdlm.
#include <cstdarg>
#include <cstdio>

#include "cdlm_internal.h"

cdlm_status cdlm_fail(cdlm_ctx *ctx, cdlm_status st, const char *fmt, ...) {
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return st;
}

extern "C" cdlm_status cdlm_create(cdlm_ctx **out, int32_t device) {
    if (!out) return CDLM_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return CDLM_ERR_HIP;
    }
    if (device < 0 || device >= ndev) return CDLM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return CDLM_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CDLM_ERR_HIP;
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) {
        // the kernels are built for gfx950 only
        return CDLM_ERR_UNSUPPORTED;
    }
    cdlm_ctx *c = new cdlm_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CDLM_ERR_HIP;
    }
    if (cdlm_pwm_scratch_init(c) != CDLM_OK) {   // this device's copy of the fixed scan scratch
        (void)hipStreamDestroy(c->stream);
        delete c;
        return CDLM_ERR_HIP;
    }
    *out = c;
    return CDLM_OK;
}

extern "C" void cdlm_destroy(cdlm_ctx *ctx) {
    if (!ctx) return;
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

extern "C" const char *cdlm_last_error(const cdlm_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

extern "C" const char *cdlm_version(void) { return "cdlm 0.1.0 (gfx950)"; }

extern "C" cdlm_status cdlm_mem_info(cdlm_ctx *ctx, size_t *free_bytes, size_t *total_bytes) {
    if (!ctx) return CDLM_ERR_ARG;
    size_t f = 0, t = 0;
    CDLM_HIP(ctx, hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return CDLM_OK;
}

namespace {
// Each thread moves 4 float4 (loads issued before stores).
__global__ __launch_bounds__(256) void k_copy4(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n4) {
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const size_t i = base + 256 * u;
        if (i < n4) v[u] = a[i];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const size_t i = base + 256 * u;
        if (i < n4) b[i] = v[u];
    }
}
}  // namespace

extern "C" cdlm_status cdlm_copy_bench(cdlm_ctx *ctx, size_t bytes, int32_t reps, double *ms_per_copy) {
    if (!ctx || reps <= 0 || bytes < 16) return CDLM_ERR_ARG;
    const size_t n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    CDLM_HIP(ctx, hipMalloc(&a, n4 * 16));
    if (hipMalloc(&b, n4 * 16) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(a);
        return cdlm_fail(ctx, CDLM_ERR_OOM, "copy buffers");
    }
    hipEvent_t e0, e1;
    cdlm_status st = CDLM_OK;
    if (hipMemsetAsync(a, 0, n4 * 16, ctx->stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        st = cdlm_fail(ctx, CDLM_ERR_HIP, "copy setup failed");
    } else {
        const unsigned blocks = (unsigned)((n4 + 1023) / 1024);
        hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(256), 0, ctx->stream, a, b, n4);  // warm-up
        (void)hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(256), 0, ctx->stream, a, b, n4);
        (void)hipEventRecord(e1, ctx->stream);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            st = cdlm_fail(ctx, CDLM_ERR_HIP, "copy timing failed");
        else if (ms_per_copy) *ms_per_copy = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}
