// Classic RK4 over fused stage kernels, one step captured into a hipGraph.
//
// Reference: TimeDependentPDE::advanceRungeKutta4 (src/pdes/time_dependent_pde.cpp:172-304):
//   k1 = f(u); s = u + 0.5 dt k1; k2 = f(s); s = u + 0.5 dt k2; k3 = f(s); s = u + dt k3;
//   k4 = f(s); u += dt/6 (k1 + 2 k2 + 2 k3 + k4)
// and the OpenCL variant's fusion of "f then stage input" into one kernel
// (src/cl_kernels/rk4_thread_per_stencil.cl:13-86) plus a final combine (89-138).
//
// Here each stage is ONE kernel: gather-SpMM of all operators for FPT fields of a row, the
// right-hand side f = sum_o c_o (D_o X), and the stage update, with a running FP64
// accumulator instead of four k arrays:
//   stage 1: acc = f          X' = u + (0.5 dt) f
//   stage 2: acc = acc + 2 f  X' = u + (0.5 dt) f
//   stage 3: acc = acc + 2 f  X' = u + dt f
//   stage 4: u = u + (dt/6) (acc + f)       [X' = u in storage precision]
// The additions and products are kept unfused (dadd_nc / dmul_nc, written under
// `#pragma clang fp contract(off)`) so they round exactly where the reference's loops round. State u and acc are FP64; X (the gathered stage field)
// is FP64 or FP32 storage. In the FP64 path stage 1 gathers u itself.
#include <cstring>

#include "rbfx_ell_common.h"

struct rbfx_rk4 {
    rbfx_ctx* ctx = nullptr;
    rbfx_ell* ell = nullptr;
    int F = 0, field_storage = RBFX_F64, accum = RBFX_F64, use_graph = 1;
    const double* coef[4] = {nullptr, nullptr, nullptr, nullptr};
    double* u = nullptr;
    double* acc = nullptr;
    void* xa = nullptr;
    void* xb = nullptr;
    bool have_state = false;
    EllTune tune;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double graph_dt = 0.0;
    int graph_layout = -1;  // operator-set layout version the graph was captured with
};

namespace {

struct StageArgs {
    const int32_t* idx_t;
    const void* w_t;
    const int32_t* order;
    int n_rows, width, F;
    const void* xin;
    double* u;
    double* acc;
    void* xout;
    const double* coef[4];
    double hdt, dt, dt6;
    int stage;
    int n_tiles, nb;
    // LDS staging of the operator set (rk4_stage_lds_kernel only)
    const int32_t* loc_pos;
    const int32_t* loc_off;
    const int32_t* loc_nodes;
    const uint16_t* loc_idx;
    int lt;
};

// Right-hand side f = sum_o c_o (D_o X) and the stage update of one thread's FPT fields.
template <typename UT, typename AT, int NOPS, int FPT>
__device__ __forceinline__ void stage_update(const StageArgs& a, int row, int fofs, const AT (&d)[NOPS][FPT]);

template <typename WT, typename UT, typename AT, int NOPS, int FPT, int UB, int SG, int PD>
__global__ void __launch_bounds__(256) rk4_stage_kernel(StageArgs a) {
    const int G = a.F / FPT;
    const int R = blockDim.x / G;
    int tile = xcd_tile(blockIdx.x, a.nb);
    if (tile >= a.n_tiles) return;
    int t = threadIdx.x;
    int rl = t / G;
    if (rl >= R) return;
    const int p = tile * R + rl;  // processing position (operator layout order)
    if (p >= a.n_rows) return;
    const int row = ell_row_at(a.order, p);  // node whose state this thread advances
    int fofs = (t - rl * G) * FPT;
    AT d[NOPS][FPT];
    if constexpr (SG == 0)
        ell_gather_dot<WT, UT, AT, NOPS, FPT, UB>(a.idx_t, (const WT*)a.w_t, (const UT*)a.xin, p, a.width, a.F, fofs,
                                                  d);
    else if constexpr (PD == 0)  // the launcher guarantees G == SG
        ell_gather_dot_shared<WT, UT, AT, NOPS, FPT, UB, SG>(a.idx_t, (const WT*)a.w_t, (const UT*)a.xin, p, a.width,
                                                             a.F, fofs, t - rl * G, d);
    else
        ell_gather_dot_pipe<WT, UT, AT, NOPS, FPT, UB, SG, PD>(a.idx_t, (const WT*)a.w_t, (const UT*)a.xin, p,
                                                               a.width, a.F, fofs, t - rl * G, d);
    stage_update<UT, AT, NOPS, FPT>(a, row, fofs, d);
}

// LDS-staged variant: block = one tile of a.lt positions (a.lt * G threads, up to 512: 128-row
// tiles with 4 lanes per row); the tile's distinct neighbour records are copied into LDS
// first, then every row gathers from there.
// MX: FP32 multiply-adds on the matrix cores (16 fields, 4 lanes per row).
// The FP32-accumulation instances with 4 fields per thread (the 128-row tiles' shapes) are
// compiled for six waves per SIMD (at most 80 VGPRs):
// three 512-thread blocks per CU, as many as 48 KB of staged records allow; at the 92 VGPRs the
// compiler chose unconstrained, only two fitted (FP32 stage at N = 1M, n = 50: 271 -> 266 us).
template <typename WT, typename UT, typename AT, int NOPS, int FPT, int UB, int SG, int PD, bool MX = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu((sizeof(UT) == 4 && sizeof(AT) == 4 && FPT == 4) ? 6 : 1))) rk4_stage_lds_kernel(StageArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    UT* lds = reinterpret_cast<UT*>(smem);
    const int tile = xcd_tile(blockIdx.x, a.nb);
    if (tile >= a.n_tiles) return;  // uniform over the block
    const int t = threadIdx.x;
    const int k0 = a.loc_off[tile];
    const int rl = t / SG;
    const int p = a.loc_pos[tile] + rl;
    const bool live = p < a.loc_pos[tile + 1];  // tiles hold <= lt rows
    if constexpr (MX) {
        // the first index and weight batches are requested before the staging, which does not
        // need them, so their latency overlaps the record loads (277 -> 271 us)
        LdsMfmaGather<NOPS, UB, PD> g;
        if (live) g.start(a.loc_idx, (const float*)a.w_t, p, a.width, t - rl * SG);
        ell_stage_tile<UT>((const UT*)a.xin, a.F, a.loc_nodes + k0, a.loc_off[tile + 1] - k0, lds);
        __syncthreads();
        if (!live) return;  // after the only barrier
        const int row = ell_row_at(a.order, p);
        const int fofs = (t - rl * SG) * FPT;
        AT d[NOPS][FPT];
        g.finish(lds, fofs, d);
        stage_update<UT, AT, NOPS, FPT>(a, row, fofs, d);
    } else {
        ell_stage_tile<UT>((const UT*)a.xin, a.F, a.loc_nodes + k0, a.loc_off[tile + 1] - k0, lds);
        __syncthreads();
        if (!live) return;  // after the only barrier
        const int row = ell_row_at(a.order, p);
        const int fofs = (t - rl * SG) * FPT;
        AT d[NOPS][FPT];
        ell_gather_dot_lds<WT, UT, AT, NOPS, FPT, UB, SG, PD>(a.loc_idx, (const WT*)a.w_t, lds, p, a.width, fofs,
                                                              t - rl * SG, d);
        stage_update<UT, AT, NOPS, FPT>(a, row, fofs, d);
    }
}

// The state streams (u, acc) are read and written once per stage: their loads and stores
// carry the streaming hint so they do not evict the neighbour records that the tiles around
// this one gather from L2 (measured: FP32 stage 310 -> 282 us at N = 1M, n = 50, 16 fields;
// FP64 unchanged). The stage output X_{s+1} is stored normally: the next stage gathers it.
template <typename UT, typename AT, int NOPS, int FPT>
__device__ __forceinline__ void stage_update(const StageArgs& a, int row, int fofs, const AT (&d)[NOPS][FPT]) {
    double c[NOPS];
#pragma unroll
    for (int o = 0; o < NOPS; o++) c[o] = a.coef[o] ? a.coef[o][row] : 1.0;
    double fv[FPT];
#pragma unroll
    for (int f = 0; f < FPT; f++) {
        double s = 0.0;
#pragma unroll
        for (int o = 0; o < NOPS; o++) {
            double dv = (double)d[o][f];
            s = a.coef[o] ? dadd_nc(s, dmul_nc(c[o], dv)) : dadd_nc(s, dv);
        }
        fv[f] = s;
    }
    size_t e = (size_t)row * a.F + fofs;
    double un[FPT];
    load_pack_nt<double, FPT>(a.u + e, un);
    double xo[FPT];
    if (a.stage == 1) {
#pragma unroll
        for (int f = 0; f < FPT; f++) xo[f] = dadd_nc(un[f], dmul_nc(a.hdt, fv[f]));
        store_pack_nt<double, FPT>(a.acc + e, fv);
    } else if (a.stage < 4) {
        double ac[FPT];
        load_pack_nt<double, FPT>(a.acc + e, ac);
        double h = (a.stage == 2) ? a.hdt : a.dt;
#pragma unroll
        for (int f = 0; f < FPT; f++) {
            ac[f] = dadd_nc(ac[f], 2.0 * fv[f]);
            xo[f] = dadd_nc(un[f], dmul_nc(h, fv[f]));
        }
        store_pack_nt<double, FPT>(a.acc + e, ac);
    } else {
        double ac[FPT];
        load_pack_nt<double, FPT>(a.acc + e, ac);
#pragma unroll
        for (int f = 0; f < FPT; f++) xo[f] = dadd_nc(un[f], dmul_nc(a.dt6, dadd_nc(ac[f], fv[f])));
        store_pack_nt<double, FPT>(a.u + e, xo);
    }
    if (a.xout) {
        UT xs[FPT];
#pragma unroll
        for (int f = 0; f < FPT; f++) xs[f] = (UT)xo[f];
        store_pack<UT, FPT>((UT*)a.xout + e, xs);
    }
}

__global__ void to_f32_kernel(const double* __restrict__ in, float* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

template <typename WT, typename UT, typename AT, int NOPS>
int launch_stage_fpt(rbfx_rk4* rk, StageArgs& a, int block) {
    hipStream_t s = rk->ctx->stream;
    int nb = a.nb;
    const int fpt = rk->tune.fpt, ub = rk->tune.ub;
    const int sg = rk->tune.share ? a.F / fpt : 0;  // validated by the tuner: 2 or 4, dividing ub
    if (rk->tune.share == 4) {
        const size_t lds = (size_t)rk->ell->k_max * a.F * sizeof(UT);
#define RBFX_LDS_CASE(FPT_, UB_, SG_)                                                                   \
    if (fpt == FPT_ && ub == UB_ && sg == SG_) {                                                        \
        rk4_stage_lds_kernel<WT, UT, AT, NOPS, FPT_, UB_, SG_, 2><<<nb, block, lds, s>>>(a);            \
        RBFX_LAUNCH_CHECK(rk->ctx, "rk4_stage_lds_kernel");                                            \
        return RBFX_OK;                                                                                 \
    }
        if constexpr (NOPS >= 3) {
            RBFX_LDS_CASE(4, 8, 4) RBFX_LDS_CASE(4, 4, 4) RBFX_LDS_CASE(8, 8, 2) RBFX_LDS_CASE(8, 4, 2)
        }
#undef RBFX_LDS_CASE
        return rbfx_fail(rk->ctx, RBFX_E_INVALID, "rk4: no LDS-staged kernel for fields/thread %d, batch %d, group %d",
                         fpt, ub, sg);
    }
    if (rk->tune.share == 5) {  // LDS-staged, FP32 multiply-adds on the matrix cores
        const size_t lds = (size_t)rk->ell->k_max * a.F * sizeof(UT);
        if constexpr (NOPS >= 3 && std::is_same<WT, float>::value && std::is_same<AT, float>::value) {
            if (fpt == 4 && sg == 4 && (ub == 4 || ub == 8)) {
                if (ub == 4) rk4_stage_lds_kernel<WT, UT, AT, NOPS, 4, 4, 4, 2, true><<<nb, block, lds, s>>>(a);
                else rk4_stage_lds_kernel<WT, UT, AT, NOPS, 4, 8, 4, 2, true><<<nb, block, lds, s>>>(a);
                RBFX_LAUNCH_CHECK(rk->ctx, "rk4_stage_lds_kernel");
                return RBFX_OK;
            }
        }
        return rbfx_fail(rk->ctx, RBFX_E_INVALID, "rk4: no matrix-core LDS-staged kernel for these types, fields/thread %d, "
                         "batch %d, group %d", fpt, ub, sg);
    }
    const int pd = rk->tune.share >= 2 ? rk->tune.share : 0;  // prefetch depth of the pipelined core
#define RBFX_STAGE_CASE(FPT_, UB_, SG_, PD_)                                               \
    if (fpt == FPT_ && ub == UB_ && sg == SG_ && pd == PD_) {                              \
        rk4_stage_kernel<WT, UT, AT, NOPS, FPT_, UB_, SG_, PD_><<<nb, block, 0, s>>>(a); \
        RBFX_LAUNCH_CHECK(rk->ctx, "rk4_stage_kernel");                                   \
        return RBFX_OK;                                                                    \
    }
    RBFX_STAGE_CASE(1, 4, 0, 0) RBFX_STAGE_CASE(2, 4, 0, 0) RBFX_STAGE_CASE(4, 4, 0, 0) RBFX_STAGE_CASE(8, 4, 0, 0)
    RBFX_STAGE_CASE(1, 8, 0, 0) RBFX_STAGE_CASE(2, 8, 0, 0) RBFX_STAGE_CASE(4, 8, 0, 0) RBFX_STAGE_CASE(8, 8, 0, 0)
    RBFX_STAGE_CASE(2, 4, 2, 0) RBFX_STAGE_CASE(4, 4, 2, 0) RBFX_STAGE_CASE(8, 4, 2, 0)
    RBFX_STAGE_CASE(2, 8, 2, 0) RBFX_STAGE_CASE(4, 8, 2, 0) RBFX_STAGE_CASE(8, 8, 2, 0)
    RBFX_STAGE_CASE(2, 4, 4, 0) RBFX_STAGE_CASE(4, 4, 4, 0) RBFX_STAGE_CASE(8, 4, 4, 0)
    RBFX_STAGE_CASE(2, 8, 4, 0) RBFX_STAGE_CASE(4, 8, 4, 0) RBFX_STAGE_CASE(8, 8, 4, 0)
    if constexpr (NOPS >= 3) {  // pipelined variants for the multi-operator sets
        RBFX_STAGE_CASE(4, 4, 4, 2) RBFX_STAGE_CASE(4, 8, 4, 2) RBFX_STAGE_CASE(4, 4, 4, 3) RBFX_STAGE_CASE(4, 8, 4, 3)
        RBFX_STAGE_CASE(8, 4, 2, 2) RBFX_STAGE_CASE(8, 8, 2, 2) RBFX_STAGE_CASE(8, 4, 2, 3) RBFX_STAGE_CASE(8, 8, 2, 3)
        RBFX_STAGE_CASE(4, 4, 2, 2) RBFX_STAGE_CASE(4, 8, 2, 2) RBFX_STAGE_CASE(2, 4, 4, 2) RBFX_STAGE_CASE(2, 8, 4, 2)
    }
#undef RBFX_STAGE_CASE
    return rbfx_fail(rk->ctx, RBFX_E_INVALID,
                     "rk4: unsupported shape (fields/thread %d, batch %d, shared group %d, prefetch %d)", fpt, ub, sg,
                     pd);
}

template <typename WT, typename UT, typename AT>
int launch_stage_ops(rbfx_rk4* rk, StageArgs& a, int block) {
    switch (rk->ell->n_ops) {
        case 1: return launch_stage_fpt<WT, UT, AT, 1>(rk, a, block);
        case 2: return launch_stage_fpt<WT, UT, AT, 2>(rk, a, block);
        case 3: return launch_stage_fpt<WT, UT, AT, 3>(rk, a, block);
        case 4: return launch_stage_fpt<WT, UT, AT, 4>(rk, a, block);
        default: return rbfx_fail(rk->ctx, RBFX_E_TOO_LARGE, "rk4: 1..4 operators supported");
    }
}

int launch_stage(rbfx_rk4* rk, int stage, double dt) {
    rbfx_ell* e = rk->ell;
    StageArgs a;
    std::memset(&a, 0, sizeof(a));
    a.idx_t = e->idx_t;
    a.w_t = e->w_t;
    a.order = e->order;
    a.n_rows = e->n_rows;
    a.width = e->width;
    a.F = rk->F;
    for (int o = 0; o < 4; o++) a.coef[o] = rk->coef[o];
    a.hdt = 0.5 * dt;
    a.dt = dt;
    a.dt6 = dt / 6.;
    a.stage = stage;
    a.u = rk->u;
    a.acc = rk->acc;
    bool f64 = rk->field_storage == RBFX_F64;
    // buffers: FP64: u -> xa -> xb -> xa -> u ; FP32: xb -> xa -> xb -> xa -> (u, xb)
    switch (stage) {
        case 1: a.xin = f64 ? (const void*)rk->u : rk->xb; a.xout = rk->xa; break;
        case 2: a.xin = rk->xa; a.xout = rk->xb; break;
        case 3: a.xin = rk->xb; a.xout = rk->xa; break;
        default: a.xin = rk->xa; a.xout = f64 ? nullptr : rk->xb; break;
    }
    int block = rk->tune.block;
    int G = rk->F / rk->tune.fpt;
    const bool staged = rk->tune.share == 4 || rk->tune.share == 5;
    if (staged) {  // LDS-staged: one block per tile of lt positions
        if (e->lt <= 0) return rbfx_fail(rk->ctx, RBFX_E_STATE, "rk4: LDS-staged shape without rbfx_ell_build_local");
        block = e->lt * G;
        a.loc_pos = e->loc_pos;
        a.loc_off = e->loc_off;
        a.loc_nodes = e->loc_nodes;
        a.loc_idx = e->loc_idx;
        a.lt = e->lt;
    }
    int R = block / G;
    a.n_tiles = staged ? e->n_ltiles : (e->n_rows + R - 1) / R;
    a.nb = (a.n_tiles + 7) / 8 * 8;
    if (e->storage == RBFX_F64) return launch_stage_ops<double, double, double>(rk, a, block);
    if (rk->accum == RBFX_F64) return launch_stage_ops<float, float, double>(rk, a, block);
    return launch_stage_ops<float, float, float>(rk, a, block);
}

int capture_step(rbfx_rk4* rk, double dt) {
    rbfx_ctx* ctx = rk->ctx;
    if (rk->exec) { hipGraphExecDestroy(rk->exec); rk->exec = nullptr; }
    if (rk->graph) { hipGraphDestroy(rk->graph); rk->graph = nullptr; }
    RBFX_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = RBFX_OK;
    for (int s = 1; s <= 4 && rc == RBFX_OK; s++) rc = launch_stage(rk, s, dt);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc != RBFX_OK) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return rbfx_fail(ctx, RBFX_E_HIP, "rk4: graph capture failed: %s", hipGetErrorString(e));
    rk->graph = g;
    RBFX_HIP_CHECK(ctx, hipGraphInstantiate(&rk->exec, rk->graph, nullptr, nullptr, 0));
    rk->graph_dt = dt;
    rk->graph_layout = rk->ell->layout_version;
    return RBFX_OK;
}

}  // namespace

extern "C" {

int rbfx_rk4_create(rbfx_ctx* ctx, const rbfx_rk4_desc* d, rbfx_rk4** out) {
    if (!ctx || !d || !out) return RBFX_E_INVALID;
    *out = nullptr;
    rbfx_ell* e = d->ell;
    if (!e || e->ctx != ctx) return rbfx_fail(ctx, RBFX_E_INVALID, "rk4_create: operator set missing or from another context");
    if (e->n_rows != e->n_cols) return rbfx_fail(ctx, RBFX_E_INVALID, "rk4_create: operators must be square (rows %d cols %d)", e->n_rows, e->n_cols);
    if (e->n_ops > 4) return rbfx_fail(ctx, RBFX_E_TOO_LARGE, "rk4_create: 1..4 operators supported, got %d", e->n_ops);
    if (d->n_fields <= 0) return rbfx_fail(ctx, RBFX_E_INVALID, "rk4_create: n_fields %d", d->n_fields);
    if (d->field_storage != RBFX_F64 && d->field_storage != RBFX_F32) return rbfx_fail(ctx, RBFX_E_INVALID, "rk4_create: bad field storage");
    if (d->accum != RBFX_F64 && d->accum != RBFX_F32) return rbfx_fail(ctx, RBFX_E_INVALID, "rk4_create: bad accumulation");
    if (e->storage == RBFX_F64 && (d->field_storage != RBFX_F64 || d->accum != RBFX_F64))
        return rbfx_fail(ctx, RBFX_E_INVALID, "rk4_create: FP64 operators need FP64 fields and accumulation");
    if (e->storage == RBFX_F32 && d->field_storage != RBFX_F32)
        return rbfx_fail(ctx, RBFX_E_INVALID, "rk4_create: FP32 operators need FP32 stage fields");
    RBFX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    rbfx_rk4* rk = new rbfx_rk4();
    rk->ctx = ctx;
    rk->ell = e;
    rk->F = d->n_fields;
    rk->field_storage = d->field_storage;
    rk->accum = d->accum;
    rk->use_graph = d->use_graph;
    for (int o = 0; o < e->n_ops; o++) rk->coef[o] = d->d_coef[o];
    int rc = rbfx_ell_pick(e, rk->F, rk->field_storage, rk->accum, 1, &rk->tune);
    if (rc) { delete rk; return rc; }
    size_t nf = (size_t)e->n_rows * rk->F;
    size_t xs = nf * (rk->field_storage == RBFX_F64 ? sizeof(double) : sizeof(float));
    hipError_t err;
    if ((err = hipMalloc(&rk->u, nf * sizeof(double))) != hipSuccess || (err = hipMalloc(&rk->acc, nf * sizeof(double))) != hipSuccess ||
        (err = hipMalloc(&rk->xa, xs)) != hipSuccess || (err = hipMalloc(&rk->xb, xs)) != hipSuccess) {
        rbfx_rk4_destroy(rk);
        return rbfx_fail(ctx, RBFX_E_NOMEM, "rk4_create: allocation failed: %s", hipGetErrorString(err));
    }
    *out = rk;
    return RBFX_OK;
}

void rbfx_rk4_destroy(rbfx_rk4* rk) {
    if (!rk) return;
    hipSetDevice(rk->ctx->device);
    hipStreamSynchronize(rk->ctx->stream);
    if (rk->exec) hipGraphExecDestroy(rk->exec);
    if (rk->graph) hipGraphDestroy(rk->graph);
    hipFree(rk->u);
    hipFree(rk->acc);
    hipFree(rk->xa);
    hipFree(rk->xb);
    delete rk;
}

int rbfx_rk4_set_state(rbfx_rk4* rk, const double* d_u) {
    if (!rk || !d_u) return RBFX_E_INVALID;
    rbfx_ctx* ctx = rk->ctx;
    size_t nf = (size_t)rk->ell->n_rows * rk->F;
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(rk->u, d_u, nf * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if (rk->field_storage == RBFX_F32) {
        to_f32_kernel<<<(unsigned)((nf + 255) / 256), 256, 0, ctx->stream>>>(rk->u, (float*)rk->xb, nf);
        RBFX_LAUNCH_CHECK(ctx, "to_f32_kernel");
    }
    rk->have_state = true;
    return RBFX_OK;
}

int rbfx_rk4_get_state(rbfx_rk4* rk, double* d_u) {
    if (!rk || !d_u) return RBFX_E_INVALID;
    if (!rk->have_state) return rbfx_fail(rk->ctx, RBFX_E_STATE, "rk4_get_state: no state set");
    size_t nf = (size_t)rk->ell->n_rows * rk->F;
    RBFX_HIP_CHECK(rk->ctx, hipMemcpyAsync(d_u, rk->u, nf * sizeof(double), hipMemcpyDeviceToDevice, rk->ctx->stream));
    return RBFX_OK;
}

int rbfx_rk4_step(rbfx_rk4* rk, double dt, int32_t n_steps) {
    if (!rk || n_steps < 0) return RBFX_E_INVALID;
    if (!rk->have_state) return rbfx_fail(rk->ctx, RBFX_E_STATE, "rk4_step: call rbfx_rk4_set_state first");
    RBFX_HIP_CHECK(rk->ctx, hipSetDevice(rk->ctx->device));
    EllTune t;
    int prc = rbfx_ell_pick(rk->ell, rk->F, rk->field_storage, rk->accum, 1, &t);
    if (prc) return prc;
    bool shape_changed = (t.fpt != rk->tune.fpt || t.block != rk->tune.block || t.ub != rk->tune.ub ||
                          t.share != rk->tune.share);
    rk->tune = t;
    if (rk->use_graph) {
        // the captured launches hold the operator set's buffers: recapture after a re-layout
        if (!rk->exec || rk->graph_dt != dt || shape_changed || rk->graph_layout != rk->ell->layout_version) {
            int rc = capture_step(rk, dt);
            if (rc) return rc;
        }
        for (int i = 0; i < n_steps; i++) RBFX_HIP_CHECK(rk->ctx, hipGraphLaunch(rk->exec, rk->ctx->stream));
        return RBFX_OK;
    }
    for (int i = 0; i < n_steps; i++)
        for (int s = 1; s <= 4; s++) {
            int rc = launch_stage(rk, s, dt);
            if (rc) return rc;
        }
    return RBFX_OK;
}

int rbfx_rk4_stage(rbfx_rk4* rk, int32_t stage, double dt) {
    if (!rk || stage < 1 || stage > 4) return RBFX_E_INVALID;
    if (!rk->have_state) return rbfx_fail(rk->ctx, RBFX_E_STATE, "rk4_stage: call rbfx_rk4_set_state first");
    int rc = rbfx_ell_pick(rk->ell, rk->F, rk->field_storage, rk->accum, 1, &rk->tune);
    if (rc) return rc;
    return launch_stage(rk, stage, dt);
}

}  // extern "C"
