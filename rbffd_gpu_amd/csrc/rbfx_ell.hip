// ELL operator sets: device layout build, SpMV and fused multi-operator multi-field SpMM.
//
// Replaces the reference's per-DerType padded row-major buffers and one-thread-per-row
// gather-dot (src/rbffd/rbffd_cl.cpp:215-257, 306-463; src/cl_kernels/computeDeriv.cl:7-25)
// and ViennaCL's column-major ELL vec_mul (ell_matrix_source.h:12-40). Differences by
// design: one index array serves all operators, the operators' weights of a (slot,row) are
// adjacent, several fields are gathered per neighbour, and nothing crosses PCIe per call.
#include <algorithm>
#include <cstring>

#include "rbfx_ell_common.h"

namespace {

struct WPtrs {
    const double* p[16];
};

template <typename WT>
__global__ void ell_pack_kernel(const int32_t* __restrict__ idx, WPtrs w, int n_ops, int n_rows, int n_pad, int width,
                                int32_t* __restrict__ idx_t, WT* __restrict__ w_t) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)n_pad * width;
    if (tid >= total) return;
    int j = (int)(tid / n_pad);
    int row = (int)(tid % n_pad);
    size_t dst = (size_t)j * n_pad + row;
    if (row < n_rows) {
        size_t src = (size_t)row * width + j;
        idx_t[dst] = idx[src];
        for (int o = 0; o < n_ops; o++) w_t[dst * n_ops + o] = (WT)w.p[o][src];
    } else {
        idx_t[dst] = 0;
        for (int o = 0; o < n_ops; o++) w_t[dst * n_ops + o] = (WT)0;
    }
}

__global__ void ell_validate_kernel(const int32_t* __restrict__ idx, size_t total, int n_cols, int* __restrict__ bad) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= total) return;
    int v = idx[tid];
    if (v < 0 || v >= n_cols) atomicAdd(bad, 1);
}

template <typename WT, typename UT, typename AT, int NOPS, int FPT, int UB>
__global__ void __launch_bounds__(256) ell_spmm_kernel(const int32_t* __restrict__ idx_t, const WT* __restrict__ w_t,
                                                       const UT* __restrict__ u, UT* __restrict__ out, int n_rows,
                                                       int n_pad, int width, int F, int n_tiles, int nb) {
    const int G = F / FPT;
    const int R = blockDim.x / G;
    int tile = xcd_tile(blockIdx.x, nb);
    if (tile >= n_tiles) return;
    int t = threadIdx.x;
    int rl = t / G;
    if (rl >= R) return;
    int row = tile * R + rl;
    if (row >= n_rows) return;
    int fofs = (t - rl * G) * FPT;
    AT acc[NOPS][FPT];
    ell_gather_dot<WT, UT, AT, NOPS, FPT, UB>(idx_t, w_t, u, row, n_pad, width, F, fofs, acc);
#pragma unroll
    for (int o = 0; o < NOPS; o++) {
        UT r[FPT];
#pragma unroll
        for (int f = 0; f < FPT; f++) r[f] = (UT)acc[o][f];
        store_pack<UT, FPT>(out + ((size_t)o * n_rows + row) * F + fofs, r);
    }
}

// Single operator `op` out of an interleaved set of n_ops (one field).
template <typename WT, typename UT, typename AT>
__global__ void __launch_bounds__(256) ell_spmv_strided_kernel(const int32_t* __restrict__ idx_t,
                                                               const WT* __restrict__ w_t, const UT* __restrict__ u,
                                                               UT* __restrict__ out, int n_rows, int n_pad, int width,
                                                               int n_ops, int op, int n_tiles, int nb) {
    int tile = xcd_tile(blockIdx.x, nb);
    if (tile >= n_tiles) return;
    int row = tile * blockDim.x + threadIdx.x;
    if (row >= n_rows) return;
    AT acc = 0;
    for (int j = 0; j < width; j++) {
        size_t e = (size_t)j * n_pad + row;
        int nb_ = idx_t[e];
        acc = fma_acc<AT>((AT)w_t[e * n_ops + op], (AT)u[nb_], acc);
    }
    out[row] = (UT)acc;
}

inline int grid_blocks(int n_tiles) { return (n_tiles + 7) / 8 * 8; }

template <typename WT, typename UT, typename AT, int NOPS>
int launch_spmm_fpt(rbfx_ell* e, const EllTune& t, int F, const void* u, void* out) {
    const int fpt = t.fpt, block = t.block;
    int G = F / fpt;
    int R = block / G;
    if (R <= 0) return rbfx_fail(e->ctx, RBFX_E_INVALID, "spmm: %d fields with %d per thread exceed block %d", F, fpt, block);
    int n_tiles = (e->n_rows + R - 1) / R;
    int nb = grid_blocks(n_tiles);
    const int32_t* it = e->idx_t;
    const WT* wt = (const WT*)e->w_t;
    const UT* uu = (const UT*)u;
    UT* oo = (UT*)out;
    hipStream_t s = e->ctx->stream;
#define RBFX_SPMM_CASE(FPT_, UB_)                                                                                   \
    if (fpt == FPT_ && t.ub == UB_) {                                                                               \
        ell_spmm_kernel<WT, UT, AT, NOPS, FPT_, UB_><<<nb, block, 0, s>>>(it, wt, uu, oo, e->n_rows, e->n_pad,      \
                                                                           e->width, F, n_tiles, nb);               \
        RBFX_LAUNCH_CHECK(e->ctx, "ell_spmm_kernel");                                                              \
        return RBFX_OK;                                                                                             \
    }
    RBFX_SPMM_CASE(1, 4) RBFX_SPMM_CASE(2, 4) RBFX_SPMM_CASE(4, 4) RBFX_SPMM_CASE(8, 4)
    RBFX_SPMM_CASE(1, 8) RBFX_SPMM_CASE(2, 8) RBFX_SPMM_CASE(4, 8) RBFX_SPMM_CASE(8, 8)
#undef RBFX_SPMM_CASE
    return rbfx_fail(e->ctx, RBFX_E_INVALID, "spmm: unsupported shape (fields/thread %d, batch %d)", fpt, t.ub);
}

template <typename WT, typename UT, typename AT>
int launch_spmm_ops(rbfx_ell* e, const EllTune& t, int F, const void* u, void* out) {
    switch (e->n_ops) {
        case 1: return launch_spmm_fpt<WT, UT, AT, 1>(e, t, F, u, out);
        case 2: return launch_spmm_fpt<WT, UT, AT, 2>(e, t, F, u, out);
        case 3: return launch_spmm_fpt<WT, UT, AT, 3>(e, t, F, u, out);
        case 4: return launch_spmm_fpt<WT, UT, AT, 4>(e, t, F, u, out);
        default: return rbfx_fail(e->ctx, RBFX_E_TOO_LARGE, "spmm: fused application supports 1..4 operators, set has %d", e->n_ops);
    }
}

int check_types(rbfx_ell* e, int field_type, int accum) {
    if (field_type != RBFX_F64 && field_type != RBFX_F32) return rbfx_fail(e->ctx, RBFX_E_INVALID, "bad field type %d", field_type);
    if (accum != RBFX_F64 && accum != RBFX_F32) return rbfx_fail(e->ctx, RBFX_E_INVALID, "bad accumulation type %d", accum);
    if (e->storage == RBFX_F64 && (field_type != RBFX_F64 || accum != RBFX_F64))
        return rbfx_fail(e->ctx, RBFX_E_INVALID, "FP64-stored operators take FP64 fields and FP64 accumulation");
    if (e->storage == RBFX_F32 && field_type == RBFX_F64 && accum == RBFX_F32)
        return rbfx_fail(e->ctx, RBFX_E_INVALID, "FP64 fields with FP32 accumulation is not offered");
    return RBFX_OK;
}

}  // namespace

// Fields-per-thread choice for the fused kernels. A tuned entry (see rbfx_tune.hip) wins;
// otherwise: the largest of 8/4/2/1 dividing F with at most 16 FP64 accumulators per thread.
int rbfx_ell_pick(rbfx_ell* ell, int n_fields, int field_type, int accum, int kind, EllTune* t);

extern "C" {

int rbfx_ell_create(rbfx_ctx* ctx, int32_t n_rows, int32_t n_cols, int32_t width, const int32_t* d_idx, int32_t n_ops,
                    const uint32_t* dertype_bits, const double* const* d_w, int32_t storage, rbfx_ell** out) {
    if (!ctx || !out) return RBFX_E_INVALID;
    *out = nullptr;
    if (n_rows <= 0 || n_cols <= 0 || width <= 0 || !d_idx || n_ops <= 0 || n_ops > 16 || !d_w)
        return rbfx_fail(ctx, RBFX_E_INVALID, "ell_create: bad sizes (rows %d cols %d width %d ops %d)", n_rows, n_cols, width, n_ops);
    if (storage != RBFX_F64 && storage != RBFX_F32) return rbfx_fail(ctx, RBFX_E_INVALID, "ell_create: bad storage %d", storage);
    for (int o = 0; o < n_ops; o++)
        if (!d_w[o]) return rbfx_fail(ctx, RBFX_E_INVALID, "ell_create: weight pointer %d is NULL", o);
    RBFX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    rbfx_ell* e = new rbfx_ell();
    e->ctx = ctx;
    e->n_rows = n_rows;
    e->n_cols = n_cols;
    e->width = width;
    e->n_ops = n_ops;
    e->storage = storage;
    e->n_pad = (n_rows + ELL_ROW_PAD - 1) / ELL_ROW_PAD * ELL_ROW_PAD;
    for (int o = 0; o < n_ops; o++) e->types[o] = dertype_bits ? dertype_bits[o] : 0u;
    size_t slots = (size_t)e->n_pad * width;
    size_t wbytes = slots * n_ops * (storage == RBFX_F64 ? sizeof(double) : sizeof(float));
    int* d_bad = nullptr;
    hipError_t err;
    if ((err = hipMalloc(&e->idx_t, slots * sizeof(int32_t))) != hipSuccess || (err = hipMalloc(&e->w_t, wbytes)) != hipSuccess ||
        (err = hipMalloc(&d_bad, sizeof(int))) != hipSuccess) {
        hipFree(e->idx_t);
        hipFree(e->w_t);
        delete e;
        return rbfx_fail(ctx, RBFX_E_NOMEM, "ell_create: allocation failed: %s", hipGetErrorString(err));
    }
    // validate indices before anything gathers through them
    int bad = 0;
    size_t total_in = (size_t)n_rows * width;
    hipMemsetAsync(d_bad, 0, sizeof(int), ctx->stream);
    ell_validate_kernel<<<(unsigned)((total_in + 255) / 256), 256, 0, ctx->stream>>>(d_idx, total_in, n_cols, d_bad);
    hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    err = hipStreamSynchronize(ctx->stream);
    hipFree(d_bad);
    if (err != hipSuccess || bad) {
        int code = err != hipSuccess ? RBFX_E_HIP : RBFX_E_INVALID;
        rbfx_ell_destroy(e);
        if (code == RBFX_E_HIP) return rbfx_fail(ctx, code, "ell_create: validation failed: %s", hipGetErrorString(err));
        return rbfx_fail(ctx, code, "ell_create: %d stencil indices outside [0,%d)", bad, n_cols);
    }
    WPtrs wp;
    std::memset(&wp, 0, sizeof(wp));
    for (int o = 0; o < n_ops; o++) wp.p[o] = d_w[o];
    unsigned blocks = (unsigned)((slots + 255) / 256);
    if (storage == RBFX_F64)
        ell_pack_kernel<double><<<blocks, 256, 0, ctx->stream>>>(d_idx, wp, n_ops, n_rows, e->n_pad, width, e->idx_t, (double*)e->w_t);
    else
        ell_pack_kernel<float><<<blocks, 256, 0, ctx->stream>>>(d_idx, wp, n_ops, n_rows, e->n_pad, width, e->idx_t, (float*)e->w_t);
    err = hipGetLastError();
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    if (err != hipSuccess) {
        rbfx_ell_destroy(e);
        return rbfx_fail(ctx, RBFX_E_HIP, "ell_create: pack failed: %s", hipGetErrorString(err));
    }
    *out = e;
    return RBFX_OK;
}

void rbfx_ell_destroy(rbfx_ell* e) {
    if (!e) return;
    hipSetDevice(e->ctx->device);
    hipStreamSynchronize(e->ctx->stream);
    hipFree(e->idx_t);
    hipFree(e->w_t);
    delete e;
}

int rbfx_ell_info(const rbfx_ell* e, int32_t* n_rows, int32_t* width, int32_t* n_ops, int32_t* storage) {
    if (!e) return RBFX_E_INVALID;
    if (n_rows) *n_rows = e->n_rows;
    if (width) *width = e->width;
    if (n_ops) *n_ops = e->n_ops;
    if (storage) *storage = e->storage;
    return RBFX_OK;
}

int rbfx_spmm(rbfx_ell* e, int32_t n_fields, const void* d_u, void* d_out, int32_t field_type, int32_t accum) {
    if (!e) return RBFX_E_INVALID;
    if (n_fields <= 0 || !d_u || !d_out) return rbfx_fail(e->ctx, RBFX_E_INVALID, "spmm: bad arguments");
    int rc = check_types(e, field_type, accum);
    if (rc) return rc;
    EllTune t;
    rc = rbfx_ell_pick(e, n_fields, field_type, accum, 0, &t);
    if (rc) return rc;
    if (e->storage == RBFX_F64) return launch_spmm_ops<double, double, double>(e, t, n_fields, d_u, d_out);
    if (field_type == RBFX_F32 && accum == RBFX_F64) return launch_spmm_ops<float, float, double>(e, t, n_fields, d_u, d_out);
    if (field_type == RBFX_F32) return launch_spmm_ops<float, float, float>(e, t, n_fields, d_u, d_out);
    return launch_spmm_ops<float, double, double>(e, t, n_fields, d_u, d_out);
}

int rbfx_spmv(rbfx_ell* e, int32_t op, const void* d_u, void* d_out, int32_t field_type, int32_t accum) {
    if (!e) return RBFX_E_INVALID;
    if (op < 0 || op >= e->n_ops || !d_u || !d_out) return rbfx_fail(e->ctx, RBFX_E_INVALID, "spmv: bad arguments (op %d of %d)", op, e->n_ops);
    int rc = check_types(e, field_type, accum);
    if (rc) return rc;
    if (e->n_ops == 1) {
        EllTune t;
        rc = rbfx_ell_pick(e, 1, field_type, accum, 0, &t);
        if (rc) return rc;
        if (e->storage == RBFX_F64) return launch_spmm_fpt<double, double, double, 1>(e, t, 1, d_u, d_out);
        if (field_type == RBFX_F32 && accum == RBFX_F64) return launch_spmm_fpt<float, float, double, 1>(e, t, 1, d_u, d_out);
        if (field_type == RBFX_F32) return launch_spmm_fpt<float, float, float, 1>(e, t, 1, d_u, d_out);
        return launch_spmm_fpt<float, double, double, 1>(e, t, 1, d_u, d_out);
    }
    const int block = 256;
    int n_tiles = (e->n_rows + block - 1) / block;
    int nb = grid_blocks(n_tiles);
    hipStream_t s = e->ctx->stream;
    if (e->storage == RBFX_F64)
        ell_spmv_strided_kernel<double, double, double><<<nb, block, 0, s>>>(e->idx_t, (const double*)e->w_t, (const double*)d_u, (double*)d_out, e->n_rows, e->n_pad, e->width, e->n_ops, op, n_tiles, nb);
    else if (field_type == RBFX_F32 && accum == RBFX_F64)
        ell_spmv_strided_kernel<float, float, double><<<nb, block, 0, s>>>(e->idx_t, (const float*)e->w_t, (const float*)d_u, (float*)d_out, e->n_rows, e->n_pad, e->width, e->n_ops, op, n_tiles, nb);
    else if (field_type == RBFX_F32)
        ell_spmv_strided_kernel<float, float, float><<<nb, block, 0, s>>>(e->idx_t, (const float*)e->w_t, (const float*)d_u, (float*)d_out, e->n_rows, e->n_pad, e->width, e->n_ops, op, n_tiles, nb);
    else
        ell_spmv_strided_kernel<float, double, double><<<nb, block, 0, s>>>(e->idx_t, (const float*)e->w_t, (const double*)d_u, (double*)d_out, e->n_rows, e->n_pad, e->width, e->n_ops, op, n_tiles, nb);
    RBFX_LAUNCH_CHECK(e->ctx, "ell_spmv_strided_kernel");
    return RBFX_OK;
}

}  // extern "C"
