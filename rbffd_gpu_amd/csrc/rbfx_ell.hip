// ELL operator sets: device layout build, SpMV and fused multi-operator multi-field SpMM.
//
// Replaces the reference's per-DerType padded row-major buffers and one-thread-per-row
// gather-dot (src/rbffd/rbffd_cl.cpp:215-257, 306-463; src/cl_kernels/computeDeriv.cl:7-25)
// and ViennaCL's column-major ELL vec_mul (ell_matrix_source.h:12-40). Differences by
// design: one index array serves all operators, the operators' weights of a (slot,row) are
// adjacent, several fields are gathered per neighbour, and nothing crosses PCIe per call.
#include <algorithm>
#include <cstring>
#include <vector>

#include <type_traits>

#include "rbfx_ell_common.h"
#include "rbfx_host_internal.h"

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
    // destination order: tile, slot, row in tile (consecutive threads write consecutive entries)
    const size_t per_tile = (size_t)width * ELL_TILE;
    const int tile = (int)(tid / per_tile);
    const size_t rem = tid - (size_t)tile * per_tile;
    const int j = (int)(rem / ELL_TILE);
    const int row = tile * ELL_TILE + (int)(rem % ELL_TILE);
    const size_t dst = tid;
    if (row < n_rows) {
        size_t src = (size_t)row * width + j;
        idx_t[dst] = idx[src];
        for (int o = 0; o < n_ops; o++) w_t[dst * n_ops + o] = (WT)w.p[o][src];
    } else {
        idx_t[dst] = 0;
        for (int o = 0; o < n_ops; o++) w_t[dst * n_ops + o] = (WT)0;
    }
}

// FP32 storage with sum-preserving rounding, one thread per (row, operator): round each
// weight to nearest, then walk the slots in stencil order (centre first, then by distance,
// which is roughly descending magnitude) and move a weight by at most one FP32 ulp toward
// the remaining difference between the FP64 row sum and the exact sum of the FP32 values.
// Each stored weight stays within 1.5 ulp of its FP64 value, and the row sum of the stored
// weights matches the FP64 row sum to about the ulp of the smallest weights instead of the
// ulp of the largest. Derivative rows sum to zero, so the rounding no longer adds
// (sum of the rounding errors) * u[centre] to every derivative: what is left scales with
// the field's variation across the stencil, not with its magnitude.
__global__ void ell_round_f32_kernel(WPtrs w, int n_ops, int n_rows, int width, float* __restrict__ w_t) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)n_rows * n_ops) return;
    int row = (int)(tid / n_ops);
    int o = (int)(tid - (size_t)row * n_ops);
    const double* src = w.p[o] + (size_t)row * width;
    double s64 = 0.0, s32 = 0.0;
    for (int j = 0; j < width; j++) {
        double v = src[j];
        s64 += v;
        s32 += (double)(float)v;
    }
    double r = s64 - s32;
    for (int j = 0; j < width; j++) {
        float v = (float)src[j];
        if (v != 0.0f && std::isfinite(r)) {
            float a = std::fabs(v);
            double ulp = (double)nextafterf(a, INFINITY) - (double)a;
            double k = std::fmin(1.0, std::fmax(-1.0, std::rint(r / ulp)));
            if (k != 0.0 && std::isfinite(ulp)) {
                float nv = (float)((double)v + k * ulp);
                r -= (double)nv - (double)v;
                v = nv;
            }
        }
        w_t[(ell_row_base(row, width) + (size_t)j * ELL_TILE) * n_ops + o] = v;
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
                                                       const UT* __restrict__ u, UT* __restrict__ out,
                                                       const int32_t* __restrict__ order, int n_rows, int width, int F,
                                                       int n_tiles, int nb) {
    const int G = F / FPT;
    const int R = blockDim.x / G;
    int tile = xcd_tile(blockIdx.x, nb);
    if (tile >= n_tiles) return;
    int t = threadIdx.x;
    int rl = t / G;
    if (rl >= R) return;
    const int p = tile * R + rl;  // processing position (layout order)
    if (p >= n_rows) return;
    const int row = ell_row_at(order, p);
    int fofs = (t - rl * G) * FPT;
    AT acc[NOPS][FPT];
    ell_gather_dot<WT, UT, AT, NOPS, FPT, UB>(idx_t, w_t, u, p, width, F, fofs, acc);
#pragma unroll
    for (int o = 0; o < NOPS; o++) {
        UT r[FPT];
#pragma unroll
        for (int f = 0; f < FPT; f++) r[f] = (UT)acc[o][f];
        store_pack<UT, FPT>(out + ((size_t)o * n_rows + row) * F + fofs, r);
    }
}

// LDS-staged fused SpMM: one block per tile of the operator set's tile tables
// (rbfx_ell_build_local). The tile's distinct neighbour records are copied into LDS once, then
// every row gathers from there through 16-bit local indices; the summation order is that of
// ell_spmm_kernel, so the outputs are the same bit for bit. MX: FP32 multiply-adds on the
// matrix cores (16 fields on 4 lanes per row). FP32-accumulation instances with 4 fields per
// thread: six waves per SIMD (three 512-thread blocks per CU), as in rk4_stage_lds_kernel.
template <typename WT, typename UT, typename AT, int NOPS, int FPT, int UB, int SG, int PD, bool MX>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu((sizeof(UT) == 4 && sizeof(AT) == 4 && FPT == 4) ? 6 : 1))) ell_spmm_lds_kernel(const int32_t* __restrict__ loc_pos,
                                                           const int32_t* __restrict__ loc_off,
                                                           const int32_t* __restrict__ loc_nodes,
                                                           const uint16_t* __restrict__ loc_idx,
                                                           const WT* __restrict__ w_t, const UT* __restrict__ u,
                                                           UT* __restrict__ out, const int32_t* __restrict__ order,
                                                           int n_rows, int width, int F, int n_tiles, int nb) {
    extern __shared__ __align__(16) unsigned char smem[];
    UT* lds = reinterpret_cast<UT*>(smem);
    const int tile = xcd_tile(blockIdx.x, nb);
    if (tile >= n_tiles) return;  // uniform over the block
    const int k0 = loc_off[tile];
    const int t = threadIdx.x;
    const int rl = t / SG;
    const int p = loc_pos[tile] + rl;
    const bool live = p < loc_pos[tile + 1];
    const int fofs = (t - rl * SG) * FPT;
    AT d[NOPS][FPT];
    if constexpr (MX) {
        // the first index and weight batches are requested before the staging
        LdsMfmaGather<NOPS, UB, PD> g;
        if (live) g.start(loc_idx, (const float*)w_t, p, width, t - rl * SG);
        ell_stage_tile<UT>(u, F, loc_nodes + k0, loc_off[tile + 1] - k0, lds);
        __syncthreads();
        if (!live) return;  // after the only barrier
        g.finish((const float*)lds, fofs, d);
    } else {
        ell_stage_tile<UT>(u, F, loc_nodes + k0, loc_off[tile + 1] - k0, lds);
        __syncthreads();
        if (!live) return;  // after the only barrier
        ell_gather_dot_lds<WT, UT, AT, NOPS, FPT, UB, SG, PD>(loc_idx, w_t, lds, p, width, fofs, t - rl * SG, d);
    }
    const int row = ell_row_at(order, p);
#pragma unroll
    for (int o = 0; o < NOPS; o++) {
        UT r[FPT];
#pragma unroll
        for (int f = 0; f < FPT; f++) r[f] = (UT)d[o][f];
        store_pack<UT, FPT>(out + ((size_t)o * n_rows + row) * F + fofs, r);
    }
}

// Single operator `op` out of an interleaved set of n_ops (one field).
template <typename WT, typename UT, typename AT>
__global__ void __launch_bounds__(256) ell_spmv_strided_kernel(const int32_t* __restrict__ idx_t,
                                                               const WT* __restrict__ w_t, const UT* __restrict__ u,
                                                               UT* __restrict__ out, const int32_t* __restrict__ order,
                                                               int n_rows, int width, int n_ops, int op, int n_tiles,
                                                               int nb) {
    int tile = xcd_tile(blockIdx.x, nb);
    if (tile >= n_tiles) return;
    const int p = tile * blockDim.x + threadIdx.x;
    if (p >= n_rows) return;
    const int row = ell_row_at(order, p);
    AT acc = 0;
    const size_t base = ell_row_base(p, width);
    for (int j = 0; j < width; j++) {
        size_t e = base + (size_t)j * ELL_TILE;
        int nb_ = idx_t[e];
        acc = fma_acc<AT>((AT)w_t[e * n_ops + op], (AT)u[nb_], acc);
    }
    out[row] = (UT)acc;
}

// Row-order validation: every row must appear exactly once. count[] starts zeroed.
__global__ void ell_order_count_kernel(const int32_t* __restrict__ order, int n, int* __restrict__ count,
                                       int* __restrict__ bad) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int r = order[p];
    if (r < 0 || r >= n) atomicAdd(bad, 1);
    else atomicAdd(&count[r], 1);
}

__global__ void ell_order_check_kernel(const int* __restrict__ count, int n, int* __restrict__ bad) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n && count[r] != 1) atomicAdd(bad, 1);
}

// moved collects the positions whose row is not the position itself (0: the identity order).
__global__ void ell_order_moved_kernel(const int32_t* __restrict__ order, int n, int* __restrict__ moved) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n && order[p] != p) atomicAdd(moved, 1);
}

// old_pos[row] = layout position of row in the current layout.
__global__ void ell_old_pos_kernel(const int32_t* __restrict__ old_order, int n, int32_t* __restrict__ old_pos) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) old_pos[old_order ? old_order[p] : p] = p;
}

// New layout entry (position p, slot j) = old entry (old_pos[order[p]], j); pad positions keep
// index 0 and weight 0. One thread per destination entry, in destination order.
template <typename WT>
__global__ void ell_reorder_kernel(const int32_t* __restrict__ idx_o, const WT* __restrict__ w_o,
                                   const int32_t* __restrict__ order, const int32_t* __restrict__ old_pos, int n_rows,
                                   int n_pad, int width, int n_ops, int32_t* __restrict__ idx_n, WT* __restrict__ w_n) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)n_pad * width) return;
    const size_t per_tile = (size_t)width * ELL_TILE;
    const int tile = (int)(tid / per_tile);
    const size_t rem = tid - (size_t)tile * per_tile;
    const int j = (int)(rem / ELL_TILE);
    const int p = tile * ELL_TILE + (int)(rem % ELL_TILE);
    if (p < n_rows) {
        size_t src = ell_row_base(old_pos[order[p]], width) + (size_t)j * ELL_TILE;
        idx_n[tid] = idx_o[src];
        for (int o = 0; o < n_ops; o++) w_n[tid * n_ops + o] = w_o[src * n_ops + o];
    } else {
        idx_n[tid] = 0;
        for (int o = 0; o < n_ops; o++) w_n[tid * n_ops + o] = (WT)0;
    }
}

inline int grid_blocks(int n_tiles) { return (n_tiles + 7) / 8 * 8; }

// LDS-staged SpMM launch (share 4; share 5 with the matrix-core core): the block is the tile.
template <typename WT, typename UT, typename AT, int NOPS>
int launch_spmm_lds(rbfx_ell* e, const EllTune& t, int F, const void* u, void* out) {
    if (e->lt <= 0) return rbfx_fail(e->ctx, RBFX_E_STATE, "spmm: LDS-staged shape without rbfx_ell_build_local");
    const int fpt = t.fpt, ub = t.ub, G = F / fpt;
    const int block = e->lt * G;
    const size_t lds = (size_t)e->k_max * F * sizeof(UT);
    const int n_tiles = e->n_ltiles, nb = grid_blocks(n_tiles);
    hipStream_t s = e->ctx->stream;
    const WT* wt = (const WT*)e->w_t;
#define RBFX_SPMM_LDS(FPT_, UB_, SG_, MX_)                                                                              if (fpt == FPT_ && ub == UB_ && G == SG_) {                                                                              ell_spmm_lds_kernel<WT, UT, AT, NOPS, FPT_, UB_, SG_, 2, MX_><<<nb, block, lds, s>>>(                                    e->loc_pos, e->loc_off, e->loc_nodes, e->loc_idx, wt, (const UT*)u, (UT*)out, e->order, e->n_rows,                  e->width, F, n_tiles, nb);                                                                                       RBFX_LAUNCH_CHECK(e->ctx, "ell_spmm_lds_kernel");                                                                   return RBFX_OK;                                                                                                  }
    if constexpr (NOPS >= 3) {
        if (t.share == 5) {
            if constexpr (std::is_same<WT, float>::value && std::is_same<UT, float>::value && std::is_same<AT, float>::value) {
                RBFX_SPMM_LDS(4, 4, 4, true) RBFX_SPMM_LDS(4, 8, 4, true)
            }
        } else {
            RBFX_SPMM_LDS(4, 8, 4, false) RBFX_SPMM_LDS(4, 4, 4, false) RBFX_SPMM_LDS(8, 8, 2, false)
            RBFX_SPMM_LDS(8, 4, 2, false)
        }
    }
#undef RBFX_SPMM_LDS
    return rbfx_fail(e->ctx, RBFX_E_INVALID, "spmm: no LDS-staged kernel for fields/thread %d, batch %d, group %d, share %d",
                     fpt, ub, G, t.share);
}

template <typename WT, typename UT, typename AT, int NOPS>
int launch_spmm_fpt(rbfx_ell* e, const EllTune& t, int F, const void* u, void* out) {
    if (t.share == 4 || t.share == 5) return launch_spmm_lds<WT, UT, AT, NOPS>(e, t, F, u, out);
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
        ell_spmm_kernel<WT, UT, AT, NOPS, FPT_, UB_><<<nb, block, 0, s>>>(it, wt, uu, oo, e->order, e->n_rows,     \
                                                                           e->width, F, n_tiles, nb);               \
        RBFX_LAUNCH_CHECK(e->ctx, "ell_spmm_kernel");                                                              \
        return RBFX_OK;                                                                                             \
    }
    RBFX_SPMM_CASE(1, 4) RBFX_SPMM_CASE(2, 4) RBFX_SPMM_CASE(4, 4) RBFX_SPMM_CASE(8, 4)
    RBFX_SPMM_CASE(1, 8) RBFX_SPMM_CASE(2, 8) RBFX_SPMM_CASE(4, 8) RBFX_SPMM_CASE(8, 8)
    RBFX_SPMM_CASE(1, 16)  // one field per thread (the SpMV): 16 slots per memory round trip
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
    if (storage != RBFX_F64 && storage != RBFX_F32 && storage != RBFX_F32_RN)
        return rbfx_fail(ctx, RBFX_E_INVALID, "ell_create: bad storage %d", storage);
    const bool sum_preserving = storage == RBFX_F32;
    if (storage == RBFX_F32_RN) storage = RBFX_F32;
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
    if (err == hipSuccess && sum_preserving) {  // overwrites the real rows' nearest-rounded weights
        size_t pairs = (size_t)n_rows * n_ops;
        ell_round_f32_kernel<<<(unsigned)((pairs + 255) / 256), 256, 0, ctx->stream>>>(wp, n_ops, n_rows, width,
                                                                                        (float*)e->w_t);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    if (err != hipSuccess) {
        rbfx_ell_destroy(e);
        return rbfx_fail(ctx, RBFX_E_HIP, "ell_create: pack failed: %s", hipGetErrorString(err));
    }
    *out = e;
    return RBFX_OK;
}

static void free_local(rbfx_ell* e) {
    hipFree(e->loc_pos);
    hipFree(e->loc_off);
    hipFree(e->loc_nodes);
    hipFree(e->loc_idx);
    e->loc_pos = e->loc_off = e->loc_nodes = nullptr;
    e->loc_idx = nullptr;
    e->lt = e->k_max = e->n_ltiles = 0;
}

void rbfx_ell_destroy(rbfx_ell* e) {
    if (!e) return;
    hipSetDevice(e->ctx->device);
    hipStreamSynchronize(e->ctx->stream);
    hipFree(e->idx_t);
    hipFree(e->w_t);
    hipFree(e->order);
    free_local(e);
    delete e;
}

int rbfx_ell_build_local(rbfx_ell* e, int32_t tile_rows, int32_t max_nodes) {
    if (!e) return RBFX_E_INVALID;
    rbfx_ctx* ctx = e->ctx;
    if (tile_rows != 32 && tile_rows != 64 && tile_rows != 128)
        return rbfx_fail(ctx, RBFX_E_INVALID, "ell_build_local: tile_rows %d not 32, 64 or 128", tile_rows);
    if (max_nodes < 0 || (max_nodes > 0 && max_nodes < e->width))
        return rbfx_fail(ctx, RBFX_E_INVALID, "ell_build_local: max_nodes %d below the stencil width %d", max_nodes,
                         e->width);
    RBFX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const size_t entries = (size_t)e->n_pad * e->width;
    std::vector<int32_t> idx(entries);
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(idx.data(), e->idx_t, entries * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> pos, off, nodes;
    std::vector<uint16_t> lidx;
    const int k_max = rbfx_build_local_tables(idx.data(), e->n_rows, e->n_pad, e->width, ELL_TILE, tile_rows, max_nodes,
                                              pos, off, nodes, lidx);
    if (k_max < 0)
        return rbfx_fail(ctx, RBFX_E_TOO_LARGE, "ell_build_local: a tile of %d rows has more than 65535 distinct nodes", tile_rows);
    int lds_attr = 0;
    if (hipDeviceGetAttribute(&lds_attr, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess || lds_attr <= 0)
        lds_attr = 64 * 1024;
    int32_t *d_pos = nullptr, *d_off = nullptr, *d_nodes = nullptr;
    uint16_t* d_lidx = nullptr;
    hipError_t err;
    auto drop = [&]() {
        hipFree(d_pos);
        hipFree(d_off);
        hipFree(d_nodes);
        hipFree(d_lidx);
    };
    if ((err = hipMalloc(&d_pos, pos.size() * sizeof(int32_t))) != hipSuccess ||
        (err = hipMalloc(&d_off, off.size() * sizeof(int32_t))) != hipSuccess ||
        (err = hipMalloc(&d_nodes, std::max<size_t>(1, nodes.size()) * sizeof(int32_t))) != hipSuccess ||
        (err = hipMalloc(&d_lidx, lidx.size() * sizeof(uint16_t))) != hipSuccess) {
        drop();
        return rbfx_fail(ctx, RBFX_E_NOMEM, "ell_build_local: %s", hipGetErrorString(err));
    }
    hipMemcpyAsync(d_pos, pos.data(), pos.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(d_off, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(d_nodes, nodes.data(), nodes.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(d_lidx, lidx.data(), lidx.size() * sizeof(uint16_t), hipMemcpyHostToDevice, ctx->stream);
    if ((err = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
        drop();
        return rbfx_fail(ctx, RBFX_E_HIP, "ell_build_local: %s", hipGetErrorString(err));
    }
    free_local(e);
    e->loc_pos = d_pos;
    e->n_ltiles = (int)pos.size() - 1;
    e->loc_off = d_off;
    e->loc_nodes = d_nodes;
    e->loc_idx = d_lidx;
    e->lt = tile_rows;
    e->k_max = k_max;
    e->lds_limit = lds_attr;
    e->layout_version++;
    return RBFX_OK;
}

int rbfx_ell_local_info(const rbfx_ell* e, int32_t* tile_rows, int32_t* k_max, int32_t* n_tiles) {
    if (!e) return RBFX_E_INVALID;
    if (tile_rows) *tile_rows = e->lt;
    if (k_max) *k_max = e->k_max;
    if (n_tiles) *n_tiles = e->n_ltiles;
    return RBFX_OK;
}

int rbfx_ell_set_row_order(rbfx_ell* e, const int32_t* d_order) {
    if (!e) return RBFX_E_INVALID;
    rbfx_ctx* ctx = e->ctx;
    if (!d_order) return rbfx_fail(ctx, RBFX_E_INVALID, "ell_set_row_order: order is NULL");
    RBFX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int n = e->n_rows;
    const size_t entries = (size_t)e->n_pad * e->width;
    const size_t wsz = e->storage == RBFX_F64 ? sizeof(double) : sizeof(float);
    hipStream_t s = ctx->stream;
    int* d_count = nullptr;
    int32_t *d_old_pos = nullptr, *d_new_order = nullptr, *idx_n = nullptr;
    void* w_n = nullptr;
    hipError_t err;
    auto release = [&]() {
        hipStreamSynchronize(s);
        hipFree(d_count);
        hipFree(d_old_pos);
        hipFree(d_new_order);
        hipFree(idx_n);
        hipFree(w_n);
    };
    if ((err = hipMalloc(&d_count, sizeof(int) * ((size_t)n + 2))) != hipSuccess ||
        (err = hipMalloc(&d_old_pos, sizeof(int32_t) * (size_t)n)) != hipSuccess ||
        (err = hipMalloc(&d_new_order, sizeof(int32_t) * (size_t)n)) != hipSuccess ||
        (err = hipMalloc(&idx_n, sizeof(int32_t) * entries)) != hipSuccess ||
        (err = hipMalloc(&w_n, wsz * entries * e->n_ops)) != hipSuccess) {
        release();
        return rbfx_fail(ctx, RBFX_E_NOMEM, "ell_set_row_order: allocation failed: %s", hipGetErrorString(err));
    }
    // validate: a permutation of 0..n-1 (count[n] collects out-of-range entries and repeats;
    // count[n+1] the positions that move, 0 for the identity)
    int flags[2] = {0, 0};
    const unsigned g = (unsigned)((n + 255) / 256);
    hipMemsetAsync(d_count, 0, sizeof(int) * ((size_t)n + 2), s);
    ell_order_count_kernel<<<g, 256, 0, s>>>(d_order, n, d_count, d_count + n);
    ell_order_check_kernel<<<g, 256, 0, s>>>(d_count, n, d_count + n);
    ell_order_moved_kernel<<<g, 256, 0, s>>>(d_order, n, d_count + n + 1);
    hipMemcpyAsync(flags, d_count + n, sizeof(flags), hipMemcpyDeviceToHost, s);
    if ((err = hipStreamSynchronize(s)) != hipSuccess) {
        release();
        return rbfx_fail(ctx, RBFX_E_HIP, "ell_set_row_order: %s", hipGetErrorString(err));
    }
    const int bad = flags[0];
    if (bad) {
        release();
        return rbfx_fail(ctx, RBFX_E_INVALID, "ell_set_row_order: not a permutation of the %d rows", n);
    }
    hipMemcpyAsync(d_new_order, d_order, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, s);
    ell_old_pos_kernel<<<g, 256, 0, s>>>(e->order, n, d_old_pos);
    const unsigned ge = (unsigned)((entries + 255) / 256);
    if (e->storage == RBFX_F64)
        ell_reorder_kernel<double><<<ge, 256, 0, s>>>(e->idx_t, (const double*)e->w_t, d_new_order, d_old_pos, n, e->n_pad,
                                                      e->width, e->n_ops, idx_n, (double*)w_n);
    else
        ell_reorder_kernel<float><<<ge, 256, 0, s>>>(e->idx_t, (const float*)e->w_t, d_new_order, d_old_pos, n, e->n_pad,
                                                     e->width, e->n_ops, idx_n, (float*)w_n);
    if ((err = hipGetLastError()) != hipSuccess || (err = hipStreamSynchronize(s)) != hipSuccess) {
        release();
        return rbfx_fail(ctx, RBFX_E_HIP, "ell_set_row_order: reorder failed: %s", hipGetErrorString(err));
    }
    hipFree(e->idx_t);
    hipFree(e->w_t);
    hipFree(e->order);
    e->idx_t = idx_n;
    e->w_t = w_n;
    e->order = d_new_order;
    if (flags[1] == 0) {  // the identity: node order, which the kernels take without an order array
        hipFree(e->order);
        e->order = nullptr;
    }
    free_local(e);  // the local tables describe the previous layout
    e->layout_version++;
    idx_n = nullptr;
    w_n = nullptr;
    d_new_order = nullptr;
    release();
    return RBFX_OK;
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
        ell_spmv_strided_kernel<double, double, double><<<nb, block, 0, s>>>(e->idx_t, (const double*)e->w_t, (const double*)d_u, (double*)d_out, e->order, e->n_rows, e->width, e->n_ops, op, n_tiles, nb);
    else if (field_type == RBFX_F32 && accum == RBFX_F64)
        ell_spmv_strided_kernel<float, float, double><<<nb, block, 0, s>>>(e->idx_t, (const float*)e->w_t, (const float*)d_u, (float*)d_out, e->order, e->n_rows, e->width, e->n_ops, op, n_tiles, nb);
    else if (field_type == RBFX_F32)
        ell_spmv_strided_kernel<float, float, float><<<nb, block, 0, s>>>(e->idx_t, (const float*)e->w_t, (const float*)d_u, (float*)d_out, e->order, e->n_rows, e->width, e->n_ops, op, n_tiles, nb);
    else
        ell_spmv_strided_kernel<float, double, double><<<nb, block, 0, s>>>(e->idx_t, (const float*)e->w_t, (const double*)d_u, (double*)d_out, e->order, e->n_rows, e->width, e->n_ops, op, n_tiles, nb);
    RBFX_LAUNCH_CHECK(e->ctx, "ell_spmv_strided_kernel");
    return RBFX_OK;
}

}  // extern "C"
