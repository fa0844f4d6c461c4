// Internal declarations shared by the HIP translation units of librbfx.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "rbfx.h"

// FP64 add and multiply that round on their own. hipcc contracts a * b + c into a fused
// multiply-add by default, also across inlined calls (this HIP version's __dadd_rn and
// __dmul_rn are plain + and *); operations written under contract(off) carry no contraction
// permission, so these two stay separate wherever they are inlined.
__device__ __forceinline__ double dadd_nc(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double dmul_nc(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}

struct rbfx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int n_cu = 0;
    int clock_khz = 0;
    char name[256] = {0};
    std::string err;
};

// Records a formatted message on ctx (or the global slot) and returns code.
int rbfx_fail(rbfx_ctx* ctx, int code, const char* fmt, ...);

#define RBFX_HIP_CHECK(ctx, call)                                                             \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return rbfx_fail((ctx), RBFX_E_HIP, "%s failed: %s (%s:%d)", #call,                \
                             hipGetErrorString(e_), __FILE__, __LINE__);                      \
    } while (0)

// Checks the last launch on this thread.
#define RBFX_LAUNCH_CHECK(ctx, what)                                                          \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            return rbfx_fail((ctx), RBFX_E_HIP, "launch of %s failed: %s", (what),            \
                             hipGetErrorString(e_));                                          \
    } while (0)

static inline unsigned popcount32(uint32_t x) { return (unsigned)__builtin_popcount(x); }
