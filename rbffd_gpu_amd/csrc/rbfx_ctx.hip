// Context, error reporting, device memory and timing helpers of the C ABI.
#include <algorithm>

#include "rbfx_internal.h"

static thread_local std::string g_create_err;

// Plain streaming copy, 16 bytes per thread per iteration: the reference point for HBM
// bandwidth (and for scaling the rocprofv3 FETCH_SIZE / WRITE_SIZE counters).
typedef unsigned v4u __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) stream_copy_kernel(const v4u* __restrict__ src, v4u* __restrict__ dst, size_t n16) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = __builtin_nontemporal_load(src + i);
}

int rbfx_fail(rbfx_ctx* ctx, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    else g_create_err = buf;
    return code;
}

extern "C" {

int rbfx_create(int device_ordinal, rbfx_ctx** out) {
    if (!out) return rbfx_fail(nullptr, RBFX_E_INVALID, "rbfx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return rbfx_fail(nullptr, RBFX_E_NODEVICE, "rbfx_create: no HIP device (%s)", hipGetErrorString(e));
    if (device_ordinal < 0 || device_ordinal >= n)
        return rbfx_fail(nullptr, RBFX_E_INVALID, "rbfx_create: device %d out of range [0,%d)", device_ordinal, n);
    e = hipSetDevice(device_ordinal);
    if (e != hipSuccess) return rbfx_fail(nullptr, RBFX_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    rbfx_ctx* c = new rbfx_ctx();
    c->device = device_ordinal;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_ordinal) == hipSuccess) {
        std::snprintf(c->name, sizeof(c->name), "%s (%s)", prop.name, prop.gcnArchName);
        c->n_cu = prop.multiProcessorCount;
        c->clock_khz = prop.clockRate;
    }
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&c->t0)) != hipSuccess || (e = hipEventCreate(&c->t1)) != hipSuccess) {
        delete c;
        return rbfx_fail(nullptr, RBFX_E_HIP, "rbfx_create: %s", hipGetErrorString(e));
    }
    *out = c;
    return RBFX_OK;
}

void rbfx_destroy(rbfx_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    hipEventDestroy(ctx->t0);
    hipEventDestroy(ctx->t1);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* rbfx_last_error(const rbfx_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int rbfx_sync(rbfx_ctx* ctx) {
    if (!ctx) return RBFX_E_INVALID;
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return RBFX_OK;
}

int rbfx_device_info(rbfx_ctx* ctx, char* name, int32_t name_len, int32_t* n_cu, int32_t* clock_khz) {
    if (!ctx) return RBFX_E_INVALID;
    if (name && name_len > 0) std::snprintf(name, (size_t)name_len, "%s", ctx->name);
    if (n_cu) *n_cu = ctx->n_cu;
    if (clock_khz) *clock_khz = ctx->clock_khz;
    return RBFX_OK;
}

int rbfx_dev_alloc(rbfx_ctx* ctx, uint64_t bytes, void** d_ptr) {
    if (!ctx || !d_ptr) return RBFX_E_INVALID;
    *d_ptr = nullptr;
    if (bytes == 0) return RBFX_OK;
    hipError_t e = hipMalloc(d_ptr, bytes);
    if (e != hipSuccess) return rbfx_fail(ctx, RBFX_E_NOMEM, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    return RBFX_OK;
}

int rbfx_dev_free(rbfx_ctx* ctx, void* d_ptr) {
    if (!ctx) return RBFX_E_INVALID;
    if (!d_ptr) return RBFX_OK;
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    RBFX_HIP_CHECK(ctx, hipFree(d_ptr));
    return RBFX_OK;
}

int rbfx_h2d(rbfx_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes) {
    if (!ctx || (bytes && (!d_dst || !h_src))) return RBFX_E_INVALID;
    if (!bytes) return RBFX_OK;
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return RBFX_OK;
}

int rbfx_d2h(rbfx_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes) {
    if (!ctx || (bytes && (!h_dst || !d_src))) return RBFX_E_INVALID;
    if (!bytes) return RBFX_OK;
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return RBFX_OK;
}

int rbfx_d2d(rbfx_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes) {
    if (!ctx || (bytes && (!d_dst || !d_src))) return RBFX_E_INVALID;
    if (!bytes) return RBFX_OK;
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return RBFX_OK;
}

int rbfx_stream_copy(rbfx_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes) {
    if (!ctx || !d_dst || !d_src || (bytes & 15) || (((uintptr_t)d_dst | (uintptr_t)d_src) & 15))
        return rbfx_fail(ctx, RBFX_E_INVALID, "stream_copy: pointers and size must be 16-byte aligned");
    size_t n16 = bytes / 16;
    unsigned blocks = (unsigned)std::min<size_t>((n16 + 255) / 256, (size_t)ctx->n_cu * 16);
    if (blocks == 0) return RBFX_OK;
    stream_copy_kernel<<<blocks, 256, 0, ctx->stream>>>((const v4u*)d_src, (v4u*)d_dst, n16);
    RBFX_LAUNCH_CHECK(ctx, "stream_copy_kernel");
    return RBFX_OK;
}

int rbfx_timer_start(rbfx_ctx* ctx) {
    if (!ctx) return RBFX_E_INVALID;
    RBFX_HIP_CHECK(ctx, hipEventRecord(ctx->t0, ctx->stream));
    return RBFX_OK;
}

int rbfx_timer_stop(rbfx_ctx* ctx, double* ms) {
    if (!ctx || !ms) return RBFX_E_INVALID;
    RBFX_HIP_CHECK(ctx, hipEventRecord(ctx->t1, ctx->stream));
    RBFX_HIP_CHECK(ctx, hipEventSynchronize(ctx->t1));
    float f = 0.f;
    RBFX_HIP_CHECK(ctx, hipEventElapsedTime(&f, ctx->t0, ctx->t1));
    *ms = (double)f;
    return RBFX_OK;
}

}  // extern "C"
