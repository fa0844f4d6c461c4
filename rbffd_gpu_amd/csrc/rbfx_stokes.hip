// Steady Stokes flow on the sphere on the device: the coupled velocity-pressure operator of
// StokesSteadyPDE (src/pdes/elliptic/stokes_steady_pde.cpp:61-230), the ILU(0) preconditioner
// of its velocity block (src/precond/ilu0.hpp:157-327 with ilu0_tag(0, 3N), as in
// tests/util/viennacl_gmres_poisson/main_stokes.cpp:83) and restarted GMRES with left
// preconditioning (the ViennaCL solve the reference calls), in FP64 or as FP32 inner solves
// inside FP64 refinement.
//
// Vectors use the reference's block layout: u (rows 0..N), v (N..2N), w (2N..3N), p (3N..4N)
// and the four constraint multipliers (4N..4N+3).
//
// Operator. One thread per stencil row gathers u, v, w and p of its neighbours once and forms
// the four block rows from the four operators of an ELL set (LSFC, XSFC, YSFC, ZSFC):
//   row u: -eta L u + D_x p + lambda_u, ..., row p: D_x u + D_y v + D_z w + lambda_p,
// and a reduction forms the constraint rows (the sums of u, v, w, p).
//
// Preconditioner. ilu0_tag(0, 3N) restricts ILU(0) to the velocity block, whose three
// diagonal blocks are the same matrix -eta L (rows of one component couple only to that
// component), with identity rows elsewhere. The factor of -eta L is computed once and applied
// to the three components. The dependencies of the factorization and of both sweeps follow
// the stencil graph; rows are grouped into dependency levels on the host (a row's level is one
// more than the deepest of the rows it reads), and one workgroup runs the levels in order with
// a barrier between levels. Every row accumulates in the reference's order, without fused
// multiply-adds, so the factor and the sweeps match the CPU restatement bit for bit.
//
// GMRES. Arnoldi with classical Gram-Schmidt applied twice (each pass one multi-dot kernel over
// the basis and one multi-axpy), Givens rotations on the host. The stopping test is the
// reference's: the estimate of ||M^-1 (b - A x)|| / ||b|| below tol.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rbfx_ell_common.h"

namespace {

constexpr int FACTOR_THREADS = 1024;
constexpr int FACTOR_WAVES = FACTOR_THREADS / 64;
constexpr int SWEEP_THREADS = 512;      // a level holds a few dozen rows, up to about a hundred
constexpr int SWEEP_CH = 32;            // entries of a row on one side of the diagonal, n <= 33
constexpr int SWEEP_CH_MAX = 64;        // the same, n <= 64
constexpr int SWEEP_EPL = 2;            // entries per lane: a row takes CH / 2 lanes
__host__ __device__ constexpr int sweep_slots(int ch) { return SWEEP_THREADS / (ch / SWEEP_EPL); }
constexpr int SWEEP_LDS_LEVELS = 8192;  // level offsets kept in LDS up to this many levels
constexpr int FACTOR_MAX_WIDTH = 64;  // one lane per stencil entry in the factorization
constexpr int FACTOR_CHUNK = 8;       // L entries whose U rows a wave loads at once
constexpr int DOT_THREADS = 256;
constexpr int DOT_BLOCKS = 256;

__device__ __forceinline__ float fadd_nc(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ float fmul_nc(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double add_nc(double a, double b) { return dadd_nc(a, b); }
__device__ __forceinline__ double mul_nc(double a, double b) { return dmul_nc(a, b); }
__device__ __forceinline__ float add_nc(float a, float b) { return fadd_nc(a, b); }
__device__ __forceinline__ float mul_nc(float a, float b) { return fmul_nc(a, b); }

template <typename WT>
__device__ __forceinline__ void load_w4(const WT* p, WT (&w)[4]) {
#pragma unroll
    for (int o = 0; o < 4; o++) w[o] = p[o];
}

// y = A x for the stencil rows (constraint_rows_kernel writes the last four).
template <typename WT, typename T>
__global__ void __launch_bounds__(256) stokes_apply_kernel(const int32_t* __restrict__ idx_t, const WT* __restrict__ w_t,
                                                           const int32_t* __restrict__ order, int N, int width, T eta,
                                                           const T* __restrict__ x, T* __restrict__ y) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int row = ell_row_at(order, p);
    const size_t base = ell_row_base(p, width);
    const T* xu = x;
    const T* xv = x + N;
    const T* xw = x + 2 * (size_t)N;
    const T* xp = x + 3 * (size_t)N;
    T lu = 0, lv = 0, lw = 0, gx = 0, gy = 0, gz = 0, dv = 0;
    for (int j = 0; j < width; j++) {
        const size_t e = base + (size_t)j * ELL_TILE;
        const int nb = idx_t[e];
        WT w[4];
        load_w4<WT>(w_t + e * 4, w);
        const T a = xu[nb], b = xv[nb], c = xw[nb], q = xp[nb];
        lu += (T)w[0] * a;
        lv += (T)w[0] * b;
        lw += (T)w[0] * c;
        gx += (T)w[1] * q;
        gy += (T)w[2] * q;
        gz += (T)w[3] * q;
        dv += (T)w[1] * a + (T)w[2] * b + (T)w[3] * c;
    }
    const T* lam = x + 4 * (size_t)N;
    y[row] = -eta * lu + gx + lam[0];
    y[N + row] = -eta * lv + gy + lam[1];
    y[2 * (size_t)N + row] = -eta * lw + gz + lam[2];
    y[3 * (size_t)N + row] = dv + lam[3];
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / 64); i++) s += red[i];
    __syncthreads();
    return s;
}

// partial[k * gridDim.x + b] = sum over block b's share of V[k*ldv + i] * w[i] (w == nullptr:
// the plain sum of V[k]); blockIdx.y = k.
template <typename T>
__global__ void __launch_bounds__(DOT_THREADS) mdot_partial_kernel(const T* __restrict__ V, size_t ldv,
                                                                   const T* __restrict__ w, size_t n,
                                                                   double* __restrict__ partial) {
    __shared__ double red[DOT_THREADS / 64];
    const int k = blockIdx.y;
    const T* v = V + (size_t)k * ldv;
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += w ? (double)v[i] * (double)w[i] : (double)v[i];
    s = block_sum<double>(s, red);
    if (threadIdx.x == 0) partial[(size_t)k * gridDim.x + blockIdx.x] = s;
}

// out[k] (+)= sum_b partial[k * nb + b], in block order; one wave per k.
__global__ void mdot_final_kernel(const double* __restrict__ partial, int nb, int K, double* __restrict__ out,
                                  int accumulate) {
    const int k = blockIdx.x;
    if (k >= K) return;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 64) s += partial[(size_t)k * nb + b];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (threadIdx.x == 0) out[k] = accumulate ? out[k] + s : s;
}

// w[i] += sign * sum_k c[k] * V[k*ldv + i]
template <typename T>
__global__ void __launch_bounds__(256) maxpy_kernel(T* __restrict__ w, const T* __restrict__ V, size_t ldv, int K,
                                                    const double* __restrict__ c, double sign, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < K; k++) s += c[k] * (double)V[(size_t)k * ldv + i];
    w[i] = (T)((double)w[i] + sign * s);
}

// y = a * x + b * y (b == 0: y = a * x); x may be y
template <typename T, typename U>
__global__ void __launch_bounds__(256) axpby_kernel(T* y, const U* x, double a, double b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = (T)(a * (double)x[i] + (b == 0.0 ? 0.0 : b * (double)y[i]));
}

// Writes the constraint rows y[4N + c] from the per-block sums of x's blocks.
template <typename T>
__global__ void constraint_rows_kernel(const double* __restrict__ sums, int N, T* __restrict__ y) {
    if (threadIdx.x < 4) y[4 * (size_t)N + threadIdx.x] = (T)sums[threadIdx.x];
}

// ILU(0) of A_L (CSR, sorted columns, diagonal present) in dependency levels. One wave per row:
// lane q holds entry q of the row in LDS; for every L entry k (ascending) the multiplier
// t = w_k / U_kk is formed and the lanes that hold row k's U entries subtract t * U_kj from the
// row's entry j where the input holds a nonzero (no fill). Nothing but w depends on the order of
// the L entries, so for FACTOR_CHUNK of them at a time the wave first loads, all at once, each
// row k's U entries and finds where their columns sit in the row (or that they do not, or that
// the input holds a zero there); the ordered part then only reads LDS.
__global__ void __launch_bounds__(FACTOR_THREADS) ilu0_factor_kernel(const int32_t* __restrict__ ip,
                                                                    const int32_t* __restrict__ col,
                                                                    const int32_t* __restrict__ diag,
                                                                    const double* __restrict__ a_in,
                                                                    const int32_t* __restrict__ rows,
                                                                    const int32_t* __restrict__ off, int n_lev,
                                                                    int nnz, double* __restrict__ f) {
    __shared__ double wv[FACTOR_WAVES][FACTOR_MAX_WIDTH];
    __shared__ int wc[FACTOR_WAVES][FACTOR_MAX_WIDTH];
    __shared__ double ukk[FACTOR_WAVES][FACTOR_MAX_WIDTH];
    __shared__ int upos[FACTOR_WAVES][FACTOR_CHUNK][FACTOR_MAX_WIDTH];  // position in the row, -1: none
    __shared__ double uval[FACTOR_WAVES][FACTOR_CHUNK][FACTOR_MAX_WIDTH];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int lev = 0; lev < n_lev; lev++) {
        const int r0 = off[lev], r1 = off[lev + 1];
        for (int ri = r0 + wave; ri < r1; ri += FACTOR_WAVES) {
            const int r = rows[ri];
            const int s = ip[r], len = ip[r + 1] - s, dpos = diag[r] - s;
            bool nz = false;
            if (lane < len) {
                wv[wave][lane] = a_in[s + lane];
                wc[wave][lane] = col[s + lane];
                nz = a_in[s + lane] != 0.0;
            }
            const uint64_t nzmask = __ballot(nz);  // entries the input holds (updates allowed)
            __builtin_amdgcn_wave_barrier();
            for (int q0 = 0; q0 < dpos; q0 += FACTOR_CHUNK) {
                const int nq = dpos - q0 < FACTOR_CHUNK ? dpos - q0 : FACTOR_CHUNK;
                if (lane < nq) ukk[wave][lane] = f[diag[wc[wave][q0 + lane]]];
                // all loads of the chunk first (addresses clamped into the arrays), then the searches
                int ud[FACTOR_CHUNK], ue[FACTOR_CHUNK], uj[FACTOR_CHUNK];
                double uu[FACTOR_CHUNK];
#pragma unroll
                for (int c = 0; c < FACTOR_CHUNK; c++) ud[c] = diag[wc[wave][q0 + (c < nq ? c : 0)]];
#pragma unroll
                for (int c = 0; c < FACTOR_CHUNK; c++) {
                    const int q = ud[c] + 1 + lane < nnz ? ud[c] + 1 + lane : nnz - 1;
                    ue[c] = ip[wc[wave][q0 + (c < nq ? c : 0)] + 1];
                    uj[c] = col[q];
                    uu[c] = f[q];
                }
#pragma unroll
                for (int c = 0; c < FACTOR_CHUNK; c++) {
                    int pos = -1;
                    if (c < nq && lane < ue[c] - ud[c] - 1) {
                        const int j = uj[c];
                        int lo = 0, hi = len;  // find j among the row's sorted columns
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (wc[wave][mid] < j) lo = mid + 1; else hi = mid;
                        }
                        if (lo < len && wc[wave][lo] == j && ((nzmask >> lo) & 1)) pos = lo;
                    }
                    upos[wave][c][lane] = pos;
                    uval[wave][c][lane] = uu[c];
                }
                __builtin_amdgcn_wave_barrier();
                for (int c = 0; c < nq; c++) {
                    const int qq = q0 + c;
                    const double t = wv[wave][qq] / ukk[wave][c];
                    __builtin_amdgcn_wave_barrier();
                    if (lane == 0) wv[wave][qq] = t;
                    const int pos = upos[wave][c][lane];
                    if (pos >= 0) wv[wave][pos] = dadd_nc(wv[wave][pos], -dmul_nc(t, uval[wave][c][lane]));
                    __builtin_amdgcn_wave_barrier();
                }
            }
            if (lane < len) f[s + lane] = wv[wave][lane];
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
    }
}

// One triangular sweep of the factor over the three velocity blocks, level by level. Task t
// (levels in order) is row rows[t] with its entries packed at [pk_off[t], pk_off[t+1]): the
// strictly lower entries for the forward sweep, the strictly upper ones (and the diagonal
// pk_diag[t]) for the backward sweep, columns ascending, at most CH of them. A level holds at
// most sweep_slots(CH) rows (the host splits larger ones) and each row takes a slot of CH / 2
// lanes: lane l loads entries l and l + CH / 2, gathers the three components at their
// columns and writes the products L_rq v_c[col_q] to LDS; lanes 0..2 of the slot then subtract
// them from v_c[r] one after the other in column order (entries past the row's end hold +0,
// and s - (+0) == s for every s), so each component accumulates exactly as the sequential
// sweep does, and finally (UPPER) divide by U_rr.
//
// Only the gathers depend on the previous level. A task's loads are spread over the two levels
// before its own: two levels ahead its row and entry range (locate), one level ahead its
// columns, values, own values and U_rr (load_entries), each issued after the current level's
// gathers and read only a level later, so the wait on memory a level pays is for its gathers.
// The level body has no loop (levels are split to fit the slots, CH covers the widest row),
// which keeps the compiler's memory-counter waits close to exact. Lanes past a row's end read
// its first entry (one address for all of them; the packed arrays carry CH zero entries past
// their end); a slot without a row reads nothing but task 0's place, which keeps those loads
// unconditional, and skips its entries, gathers and store.
template <typename T, int EPL>
struct SweepTask {
    int t;   // task index, -1: none
    int tt;  // t, or 0 for an empty slot
    int r, a, e;
    int pc[EPL];
    T pv[EPL];
    T s, dg;  // lanes 0..2: the row's own value of component l (written only by this task), U_rr
};

template <typename T, bool UPPER, bool LDS_OFF, int CH>
__global__ void __launch_bounds__(SWEEP_THREADS) ilu0_sweep_kernel(const int32_t* __restrict__ rows,
                                                                   const int32_t* __restrict__ off, int n_lev,
                                                                   const int32_t* __restrict__ pk_off,
                                                                   const int32_t* __restrict__ pk_col,
                                                                   const T* __restrict__ pk_val,
                                                                   const T* __restrict__ pk_diag, int N, T* v) {
    constexpr int EPL = SWEEP_EPL, SL = CH / EPL, NS = sweep_slots(CH);
    // a slot's three rows of products, padded so that the twelve summing lanes of a wave fall
    // in different banks and each row starts on 16 bytes
    constexpr int PS = (CH * (int)sizeof(T) + 16) / (int)sizeof(T);
    __shared__ int s_off[LDS_OFF ? SWEEP_LDS_LEVELS + 1 : 1];
    __shared__ __align__(16) T s_prod[NS][3][PS];
    if (LDS_OFF)
        for (int i = threadIdx.x; i <= n_lev; i += blockDim.x) s_off[i] = off[i];
    __syncthreads();
    const int slot = (int)threadIdx.x / SL, l = (int)threadIdx.x % SL;
    const int lc = l < 3 ? l : 2;  // lanes past 2 load a value they do not use
    T(*P)[PS] = s_prod[slot];
    auto level_start = [&](int lv) {
        lv = lv < n_lev ? lv : n_lev;
        return LDS_OFF ? s_off[lv] : off[lv];
    };
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto locate = [&](SweepTask<T, EPL>& X, int lev) {
        const int t = level_start(lev) + slot;
        const bool ok = t < level_start(lev + 1);
        X.t = ok ? t : -1;
        X.tt = ok ? t : 0;
        X.r = rows[X.tt];
        X.a = pk_off[X.tt];
        X.e = pk_off[X.tt + 1];
    };
    auto load_entries = [&](SweepTask<T, EPL>& X) {
        const int len = X.e - X.a;
        if (X.t >= 0) {
#pragma unroll
            for (int j = 0; j < EPL; j++) {
                const int k = l + j * SL;
                X.pc[j] = pk_col[X.a + (k < len ? k : 0)];
                X.pv[j] = pk_val[X.a + (k < len ? k : 0)];
            }
            X.s = v[(size_t)lc * N + X.r];
            if (UPPER) X.dg = pk_diag[X.t];
        }
    };
    // A slot without a row skips its loads and keeps the (defined) values of an earlier task in
    // the task and gather registers; nothing computed from them is stored.
    T g[3][EPL] = {};
    // X: level lev's task, loaded; Y: level lev + 1's, located; Z receives lev + 2's
    auto run = [&](SweepTask<T, EPL>& X, SweepTask<T, EPL>& Y, SweepTask<T, EPL>& Z, int lev) {
        if (X.t >= 0) {
#pragma unroll
            for (int j = 0; j < EPL; j++)
#pragma unroll
                for (int c = 0; c < 3; c++) g[c][j] = v[(size_t)c * N + X.pc[j]];
        }
        asm volatile("" ::: "memory");  // the gathers are requested before the later levels' loads
        load_entries(Y);
        locate(Z, lev + 2);
        const int len = X.t >= 0 ? X.e - X.a : 0;
#pragma unroll
        for (int j = 0; j < EPL; j++) {
            const int k = l + j * SL;
            const bool in = k < len;
#pragma unroll
            for (int c = 0; c < 3; c++) P[c][k] = in ? mul_nc(X.pv[j], g[c][j]) : T(0);
        }
        wave_sync();
        // every lane of a slot sums and stores: lanes past 2 repeat lane 2 (same value, same address)
        T s = X.s;
#pragma unroll 32
        for (int k = 0; k < CH; k++) s = add_nc(s, -P[lc][k]);
        if (UPPER) s = s / X.dg;
        if (X.t >= 0) v[(size_t)lc * N + X.r] = s;
        wave_sync();  // the slot's products are read before the next level overwrites them
        __syncthreads();
    };
    SweepTask<T, EPL> A{}, B{}, C{};
    locate(A, 0);
    load_entries(A);
    locate(B, 1);
    // levels past n_lev are empty (level_start clamps), so the loop body needs no test; the
    // compiler waits for every outstanding access at the loop head, hence the long body
    for (int lev = 0; lev < n_lev; lev += 6) {
        run(A, B, C, lev);
        run(B, C, A, lev + 1);
        run(C, A, B, lev + 2);
        run(A, B, C, lev + 3);
        run(B, C, A, lev + 4);
        run(C, A, B, lev + 5);
    }
}

// dst[i] = (T) src[map[i]]
template <typename T>
__global__ void gather_kernel(const double* __restrict__ src, const int32_t* __restrict__ map, T* __restrict__ dst,
                              int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (T)src[map[i]];
}

__global__ void to_f32_kernel(const double* __restrict__ in, float* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

inline int blocks_for(size_t n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

struct rbfx_stokes {
    rbfx_ctx* ctx = nullptr;
    rbfx_ell* e64 = nullptr;
    rbfx_ell* e32 = nullptr;
    double eta = 1.0;
    int N = 0, width = 0;
    size_t n_tot = 0;
    // factor of A_L = -eta L in CSR (sorted columns)
    int32_t *ip = nullptr, *col = nullptr, *diag = nullptr;
    double *a_in = nullptr, *f64 = nullptr;
    int32_t *lrows = nullptr, *loff = nullptr, *urows = nullptr, *uoff = nullptr;
    int n_lev_l = 0, n_lev_u = 0;
    int max_side = 0;  // most entries of a row on one side of the diagonal
    // the sweeps' packed entries (task order): offsets, columns, source index into the factor,
    // values in FP64 and FP32, and the diagonal of each backward task
    int32_t *lpk_off = nullptr, *lpk_col = nullptr, *lpk_src = nullptr;
    int32_t *upk_off = nullptr, *upk_col = nullptr, *upk_src = nullptr, *upk_dsrc = nullptr;
    double *lpk64 = nullptr, *upk64 = nullptr, *ud64 = nullptr;
    float *lpk32 = nullptr, *upk32 = nullptr, *ud32 = nullptr;
    int l_nnz = 0, u_nnz = 0;
    // GMRES workspace
    int ws_m = -1;
    void* V = nullptr;    // (m + 1) * n_tot values of the run's type (FP64 size)
    void* w = nullptr;    // n_tot
    double* r64 = nullptr;  // n_tot: outer residual of the refinement
    double* x32d = nullptr;  // n_tot: FP32 correction, scratch of the refinement
    double* partial = nullptr;  // (m + 1) * DOT_BLOCKS
    double* hbuf = nullptr;     // 2 * (m + 2): CGS coefficients, norm
    double* coef = nullptr;     // m + 1 coefficients for the solution update
};

namespace {

void stokes_free(rbfx_stokes* S) {
    for (void* p : {(void*)S->ip, (void*)S->col, (void*)S->diag, (void*)S->a_in, (void*)S->f64,
                    (void*)S->lrows, (void*)S->loff, (void*)S->urows, (void*)S->uoff, S->V, S->w, (void*)S->r64,
                    (void*)S->x32d, (void*)S->partial, (void*)S->hbuf, (void*)S->coef, (void*)S->lpk_off,
                    (void*)S->lpk_col, (void*)S->lpk_src, (void*)S->upk_off, (void*)S->upk_col, (void*)S->upk_src,
                    (void*)S->upk_dsrc, (void*)S->lpk64, (void*)S->upk64, (void*)S->ud64, (void*)S->lpk32,
                    (void*)S->upk32, (void*)S->ud32})
        if (p) hipFree(p);
}

template <typename T>
int upload(rbfx_ctx* ctx, T** d, const std::vector<T>& h) {
    if (hipMalloc((void**)d, std::max<size_t>(1, h.size()) * sizeof(T)) != hipSuccess)
        return rbfx_fail(ctx, RBFX_E_NOMEM, "stokes: device allocation of %zu bytes failed", h.size() * sizeof(T));
    if (!h.empty()) RBFX_HIP_CHECK(ctx, hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return RBFX_OK;
}

// y = A x (block layout, n_tot values) for the FP64 or FP32 operator set.
template <typename T>
int apply(rbfx_stokes* S, const T* x, T* y) {
    rbfx_ctx* ctx = S->ctx;
    hipStream_t st = ctx->stream;
    rbfx_ell* e = sizeof(T) == 8 ? S->e64 : S->e32;
    const int N = S->N;
    if (sizeof(T) == 8)
        stokes_apply_kernel<double, double><<<blocks_for(N, 256), 256, 0, st>>>(
            e->idx_t, (const double*)e->w_t, e->order, N, S->width, S->eta, (const double*)x, (double*)y);
    else
        stokes_apply_kernel<float, float><<<blocks_for(N, 256), 256, 0, st>>>(
            e->idx_t, (const float*)e->w_t, e->order, N, S->width, (float)S->eta, (const float*)x, (float*)y);
    RBFX_LAUNCH_CHECK(ctx, "stokes_apply_kernel");
    mdot_partial_kernel<T><<<dim3(DOT_BLOCKS, 4), DOT_THREADS, 0, st>>>(x, (size_t)N, nullptr, (size_t)N, S->partial);
    RBFX_LAUNCH_CHECK(ctx, "mdot_partial_kernel");
    mdot_final_kernel<<<4, 64, 0, st>>>(S->partial, DOT_BLOCKS, 4, S->hbuf, 0);
    RBFX_LAUNCH_CHECK(ctx, "mdot_final_kernel");
    constraint_rows_kernel<T><<<1, 64, 0, st>>>(S->hbuf, N, y);
    RBFX_LAUNCH_CHECK(ctx, "constraint_rows_kernel");
    return RBFX_OK;
}

template <typename T, bool UPPER>
void launch_sweep(const rbfx_stokes* S, hipStream_t st, const int32_t* rows, const int32_t* off, int n_lev,
                  const int32_t* pk_off, const int32_t* pk_col, const T* pk_val, const T* pk_diag, T* v) {
    const bool lds = n_lev <= SWEEP_LDS_LEVELS;
    const bool wide = S->max_side > SWEEP_CH;
    if (lds && !wide)
        ilu0_sweep_kernel<T, UPPER, true, SWEEP_CH><<<1, SWEEP_THREADS, 0, st>>>(rows, off, n_lev, pk_off, pk_col, pk_val,
                                                                                pk_diag, S->N, v);
    else if (lds)
        ilu0_sweep_kernel<T, UPPER, true, SWEEP_CH_MAX><<<1, SWEEP_THREADS, 0, st>>>(rows, off, n_lev, pk_off, pk_col,
                                                                                    pk_val, pk_diag, S->N, v);
    else if (!wide)
        ilu0_sweep_kernel<T, UPPER, false, SWEEP_CH><<<1, SWEEP_THREADS, 0, st>>>(rows, off, n_lev, pk_off, pk_col,
                                                                                 pk_val, pk_diag, S->N, v);
    else
        ilu0_sweep_kernel<T, UPPER, false, SWEEP_CH_MAX><<<1, SWEEP_THREADS, 0, st>>>(rows, off, n_lev, pk_off, pk_col,
                                                                                     pk_val, pk_diag, S->N, v);
}

// v = M^-1 v: both sweeps on the three velocity blocks; the other rows are identity rows.
template <typename T>
int precond(rbfx_stokes* S, T* v) {
    rbfx_ctx* ctx = S->ctx;
    hipStream_t st = ctx->stream;
    const bool d = sizeof(T) == 8;
    const T* lv = d ? (const T*)S->lpk64 : (const T*)S->lpk32;
    const T* uv = d ? (const T*)S->upk64 : (const T*)S->upk32;
    const T* ud = d ? (const T*)S->ud64 : (const T*)S->ud32;
    launch_sweep<T, false>(S, st, S->lrows, S->loff, S->n_lev_l, S->lpk_off, S->lpk_col, lv, nullptr, v);
    RBFX_LAUNCH_CHECK(ctx, "ilu0_sweep_kernel");
    launch_sweep<T, true>(S, st, S->urows, S->uoff, S->n_lev_u, S->upk_off, S->upk_col, uv, ud, v);
    RBFX_LAUNCH_CHECK(ctx, "ilu0_sweep_kernel");
    return RBFX_OK;
}

// out[0..K) = V[0..K)^T w (w == nullptr: plain sums), accumulate: out += instead.
template <typename T>
int mdot(rbfx_stokes* S, const T* V, int K, const T* w, double* out, int accumulate) {
    rbfx_ctx* ctx = S->ctx;
    mdot_partial_kernel<T><<<dim3(DOT_BLOCKS, K), DOT_THREADS, 0, ctx->stream>>>(V, S->n_tot, w, S->n_tot, S->partial);
    RBFX_LAUNCH_CHECK(ctx, "mdot_partial_kernel");
    mdot_final_kernel<<<K, 64, 0, ctx->stream>>>(S->partial, DOT_BLOCKS, K, out, accumulate);
    RBFX_LAUNCH_CHECK(ctx, "mdot_final_kernel");
    return RBFX_OK;
}

int workspace(rbfx_stokes* S, int m) {
    if (S->ws_m >= m) return RBFX_OK;
    for (void** p : {&S->V, &S->w, (void**)&S->r64, (void**)&S->x32d, (void**)&S->coef})
        if (*p) {
            hipFree(*p);
            *p = nullptr;
        }
    const size_t n = S->n_tot;
    bool ok = hipMalloc(&S->V, (size_t)(m + 1) * n * sizeof(double)) == hipSuccess &&
              hipMalloc(&S->w, n * sizeof(double)) == hipSuccess && hipMalloc(&S->r64, n * sizeof(double)) == hipSuccess &&
              hipMalloc(&S->x32d, n * sizeof(double)) == hipSuccess &&
              hipMalloc(&S->coef, (size_t)(m + 1) * sizeof(double)) == hipSuccess;
    if (!ok) {
        S->ws_m = -1;
        return rbfx_fail(S->ctx, RBFX_E_NOMEM, "stokes_gmres: workspace for Krylov dimension %d", m);
    }
    if (S->partial) hipFree(S->partial);
    if (S->hbuf) hipFree(S->hbuf);
    S->partial = nullptr;
    S->hbuf = nullptr;
    if (hipMalloc((void**)&S->partial, (size_t)(m + 4) * DOT_BLOCKS * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&S->hbuf, (size_t)2 * (m + 2) * sizeof(double)) != hipSuccess) {
        S->ws_m = -1;
        return rbfx_fail(S->ctx, RBFX_E_NOMEM, "stokes_gmres: reduction buffers");
    }
    S->ws_m = m;
    return RBFX_OK;
}

// Restarted GMRES(m) in type T on A x = b with left preconditioning, from the x given
// (zero_x: from x = 0). Cycles as ViennaCL counts them: max_iter / m, one fewer when m
// divides max_iter. The test is ||M^-1 r|| / ||b|| < tol, or with rel_prec
// ||M^-1 r|| / ||M^-1 r_0|| < tol. Returns 0 (converged), 1 (not converged) or a negative code.
template <typename T>
int gmres(rbfx_stokes* S, const T* b, T* x, double tol, int max_iter, int m, bool zero_x, bool rel_prec, int* iters,
          double* err) {
    rbfx_ctx* ctx = S->ctx;
    hipStream_t st = ctx->stream;
    const size_t n = S->n_tot;
    const int nb = blocks_for(n, 256);
    T* V = (T*)S->V;
    T* w = (T*)S->w;
    int rc;
    if (zero_x) RBFX_HIP_CHECK(ctx, hipMemsetAsync(x, 0, n * sizeof(T), st));
    if ((rc = mdot<T>(S, b, 1, b, S->hbuf, 0))) return rc;
    double nb2 = 0.0;
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(&nb2, S->hbuf, sizeof(double), hipMemcpyDeviceToHost, st));
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(st));
    double norm_b = std::sqrt(nb2);
    *err = 0.0;
    if (norm_b == 0.0) {
        RBFX_HIP_CHECK(ctx, hipMemsetAsync(x, 0, n * sizeof(T), st));
        return 0;
    }
    int cycles = max_iter / m;
    if (cycles > 0 && cycles * m == max_iter) cycles -= 1;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hv(m + 2);
    for (int cycle = 0; cycle <= cycles; cycle++) {
        // V_0 = M^-1 (b - A x) / beta
        if ((rc = apply<T>(S, x, V))) return rc;
        axpby_kernel<T, T><<<nb, 256, 0, st>>>(V, b, 1.0, -1.0, n);
        RBFX_LAUNCH_CHECK(ctx, "axpby_kernel");
        if ((rc = precond<T>(S, V))) return rc;
        if ((rc = mdot<T>(S, V, 1, V, S->hbuf, 0))) return rc;
        double b2 = 0.0;
        RBFX_HIP_CHECK(ctx, hipMemcpyAsync(&b2, S->hbuf, sizeof(double), hipMemcpyDeviceToHost, st));
        RBFX_HIP_CHECK(ctx, hipStreamSynchronize(st));
        const double beta = std::sqrt(b2);
        if (rel_prec && cycle == 0) {
            if (beta == 0.0) return 0;
            norm_b = beta;
        }
        *err = beta / norm_b;
        if (!(beta / norm_b >= tol)) return std::isfinite(beta) ? 0 : 1;
        axpby_kernel<T, T><<<nb, 256, 0, st>>>(V, V, 1.0 / beta, 0.0, n);
        RBFX_LAUNCH_CHECK(ctx, "axpby_kernel");
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k_used = 0;
        bool done = false;
        for (int k = 0; k < m; k++) {
            (*iters)++;
            if ((rc = apply<T>(S, V + (size_t)k * n, w))) return rc;
            if ((rc = precond<T>(S, w))) return rc;
            // classical Gram-Schmidt, twice: h = V^T w; w -= V h
            for (int pass = 0; pass < 2; pass++) {
                if ((rc = mdot<T>(S, V, k + 1, w, S->hbuf + pass * (m + 2), 0))) return rc;
                maxpy_kernel<T><<<nb, 256, 0, st>>>(w, V, n, k + 1, S->hbuf + pass * (m + 2), -1.0, n);
                RBFX_LAUNCH_CHECK(ctx, "maxpy_kernel");
            }
            if ((rc = mdot<T>(S, w, 1, w, S->hbuf + 2 * (m + 2) - 1, 0))) return rc;
            std::vector<double> h2(2 * (m + 2));
            RBFX_HIP_CHECK(ctx, hipMemcpyAsync(h2.data(), S->hbuf, h2.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            RBFX_HIP_CHECK(ctx, hipStreamSynchronize(st));
            double* hk = &H[(size_t)k * (m + 1)];  // column k of the Hessenberg matrix
            for (int i = 0; i <= k; i++) hk[i] = h2[i] + h2[(m + 2) + i];
            const double hn = std::sqrt(h2[2 * (m + 2) - 1]);
            hk[k + 1] = hn;
            for (int i = 0; i < k; i++) {  // earlier rotations
                const double a = hk[i], c = hk[i + 1];
                hk[i] = cs[i] * a + sn[i] * c;
                hk[i + 1] = -sn[i] * a + cs[i] * c;
            }
            const double den = std::hypot(hk[k], hk[k + 1]);
            cs[k] = den > 0 ? hk[k] / den : 1.0;
            sn[k] = den > 0 ? hk[k + 1] / den : 0.0;
            hk[k] = den;
            hk[k + 1] = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            k_used = k + 1;
            *err = std::fabs(g[k + 1]) / norm_b;
            if (!(*err >= tol) || hn == 0.0) {
                done = true;
                break;
            }
            axpby_kernel<T, T><<<nb, 256, 0, st>>>(V + (size_t)(k + 1) * n, w, 1.0 / hn, 0.0, n);
            RBFX_LAUNCH_CHECK(ctx, "axpby_kernel");
        }
        // y = R^-1 g, x += V y
        for (int i = k_used - 1; i >= 0; i--) {
            double s = g[i];
            for (int j = i + 1; j < k_used; j++) s -= H[(size_t)j * (m + 1) + i] * y[j];
            y[i] = s / H[(size_t)i * (m + 1) + i];
        }
        RBFX_HIP_CHECK(ctx, hipMemcpyAsync(S->coef, y.data(), k_used * sizeof(double), hipMemcpyHostToDevice, st));
        maxpy_kernel<T><<<nb, 256, 0, st>>>(x, V, n, k_used, S->coef, 1.0, n);
        RBFX_LAUNCH_CHECK(ctx, "maxpy_kernel");
        RBFX_HIP_CHECK(ctx, hipStreamSynchronize(st));  // y is a host array
        if (done && std::isfinite(*err)) return 0;
        if (!std::isfinite(*err)) return 1;
    }
    return 1;
}

}  // namespace

extern "C" {

int rbfx_stokes_create(rbfx_ctx* ctx, rbfx_ell* ops64, rbfx_ell* ops32, double eta, rbfx_stokes** out) {
    if (!ctx || !out) return RBFX_E_INVALID;
    *out = nullptr;
    const uint32_t want[4] = {RBFX_LSFC, RBFX_XSFC, RBFX_YSFC, RBFX_ZSFC};
    for (rbfx_ell* e : {ops64, ops32}) {
        if (!e) continue;
        if (e->ctx != ctx || e->n_ops != 4 || e->n_rows != e->n_cols || std::memcmp(e->types, want, sizeof(want)) != 0)
            return rbfx_fail(ctx, RBFX_E_INVALID,
                             "stokes_create: the operator sets must be square with the operators LSFC, XSFC, YSFC, ZSFC");
    }
    if (!ops64 || ops64->storage != RBFX_F64) return rbfx_fail(ctx, RBFX_E_INVALID, "stokes_create: ops64 must store FP64");
    if (ops32 && (ops32->storage != RBFX_F32 || ops32->n_rows != ops64->n_rows || ops32->width != ops64->width))
        return rbfx_fail(ctx, RBFX_E_INVALID, "stokes_create: ops32 must be the same operators in FP32 storage");
    const int N = ops64->n_rows, W = ops64->width;
    if (W > FACTOR_MAX_WIDTH) return rbfx_fail(ctx, RBFX_E_TOO_LARGE, "stokes_create: stencil width %d above %d", W, FACTOR_MAX_WIDTH);
    RBFX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    // the stencils and L weights back on the host, by row (layout position p holds row order[p])
    const size_t tiles = (size_t)((N + ELL_TILE - 1) / ELL_TILE) * ELL_TILE * W;
    std::vector<int32_t> idx_t(tiles), order(N);
    std::vector<double> w_t(tiles * 4);
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    RBFX_HIP_CHECK(ctx, hipMemcpy(idx_t.data(), ops64->idx_t, tiles * sizeof(int32_t), hipMemcpyDeviceToHost));
    RBFX_HIP_CHECK(ctx, hipMemcpy(w_t.data(), ops64->w_t, tiles * 4 * sizeof(double), hipMemcpyDeviceToHost));
    if (ops64->order)
        RBFX_HIP_CHECK(ctx, hipMemcpy(order.data(), ops64->order, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    else
        for (int i = 0; i < N; i++) order[i] = i;
    // CSR of A_L = -eta L with each row's columns ascending (the reference's compressed_matrix)
    std::vector<int32_t> ip(N + 1), col((size_t)N * W), dg(N);
    std::vector<double> a((size_t)N * W);
    std::vector<std::pair<int32_t, double>> ent(W);
    for (int p = 0; p < N; p++) {
        const int r = order[p];
        const size_t base = ell_row_base(p, W);
        for (int j = 0; j < W; j++) {
            const size_t e = base + (size_t)j * ELL_TILE;
            ent[j] = {idx_t[e], w_t[e * 4]};
        }
        std::sort(ent.begin(), ent.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        int d = -1;
        for (int j = 0; j < W; j++) {
            if (j && ent[j].first == ent[j - 1].first)
                return rbfx_fail(ctx, RBFX_E_INVALID, "stokes_create: stencil %d holds node %d twice", r, ent[j].first);
            col[(size_t)r * W + j] = ent[j].first;
            a[(size_t)r * W + j] = -eta * ent[j].second;
            if (ent[j].first == r) d = j;
        }
        if (d < 0) return rbfx_fail(ctx, RBFX_E_INVALID, "stokes_create: stencil %d does not contain its centre", r);
        dg[r] = r * W + d;
    }
    for (int r = 0; r <= N; r++) ip[r] = r * W;
    // dependency levels: the lower sweep and the factorization read rows k < r of row r's
    // pattern, the upper sweep rows k > r
    std::vector<int32_t> lev(N, 0), ulev(N, 0);
    int nl = 0, nu = 0;
    for (int r = 0; r < N; r++) {
        int m = -1;
        for (int q = ip[r]; q < dg[r]; q++) m = std::max(m, lev[col[q]]);
        lev[r] = m + 1;
        nl = std::max(nl, m + 2);
    }
    for (int r = N - 1; r >= 0; r--) {
        int m = -1;
        for (int q = dg[r] + 1; q < ip[r + 1]; q++) m = std::max(m, ulev[col[q]]);
        ulev[r] = m + 1;
        nu = std::max(nu, m + 2);
    }
    auto by_level = [&](const std::vector<int32_t>& L, int n_lev, std::vector<int32_t>& rows, std::vector<int32_t>& off) {
        off.assign(n_lev + 1, 0);
        for (int r = 0; r < N; r++) off[L[r] + 1]++;
        for (int l = 0; l < n_lev; l++) off[l + 1] += off[l];
        rows.assign(N, 0);
        std::vector<int32_t> pos(off.begin(), off.end() - 1);
        for (int r = 0; r < N; r++) rows[pos[L[r]]++] = r;
    };
    int max_side = 0;
    for (int r = 0; r < N; r++) max_side = std::max({max_side, dg[r] - ip[r], ip[r + 1] - dg[r] - 1});
    // a sweep runs at most sweep_slots rows per step: larger levels become several steps
    const int slots = sweep_slots(max_side > SWEEP_CH ? SWEEP_CH_MAX : SWEEP_CH);
    auto split = [&](std::vector<int32_t>& off) {
        std::vector<int32_t> o(1, 0);
        for (size_t l = 0; l + 1 < off.size(); l++) {
            for (int32_t b = off[l] + slots; b < off[l + 1]; b += slots) o.push_back(b);
            o.push_back(off[l + 1]);
        }
        off.swap(o);
    };
    std::vector<int32_t> lrows, loff, urows, uoff;
    by_level(lev, nl, lrows, loff);
    by_level(ulev, nu, urows, uoff);
    split(loff);
    split(uoff);
    nl = (int)loff.size() - 1;
    nu = (int)uoff.size() - 1;
    // the sweeps' entries packed task by task (level order), columns ascending
    std::vector<int32_t> lpo(N + 1, 0), lpc, lps, upo(N + 1, 0), upc, ups, uds(N);
    for (int t = 0; t < N; t++) {
        const int r = lrows[t];
        for (int q = ip[r]; q < dg[r]; q++) {
            lpc.push_back(col[q]);
            lps.push_back(q);
        }
        lpo[t + 1] = (int32_t)lpc.size();
    }
    for (int t = 0; t < N; t++) {
        const int r = urows[t];
        for (int q = dg[r] + 1; q < ip[r + 1]; q++) {
            upc.push_back(col[q]);
            ups.push_back(q);
        }
        upo[t + 1] = (int32_t)upc.size();
        uds[t] = dg[r];
    }
    // the sweeps read a whole chunk from any task's start: SWEEP_CH_MAX zero entries past the end
    const int l_nnz = (int)lpc.size(), u_nnz = (int)upc.size();
    lpc.resize(l_nnz + SWEEP_CH_MAX, 0);
    upc.resize(u_nnz + SWEEP_CH_MAX, 0);

    rbfx_stokes* S = new rbfx_stokes();
    S->ctx = ctx;
    S->e64 = ops64;
    S->e32 = ops32;
    S->eta = eta;
    S->N = N;
    S->width = W;
    S->n_tot = 4 * (size_t)N + 4;
    S->n_lev_l = nl;
    S->n_lev_u = nu;
    S->max_side = max_side;
    int rc = RBFX_OK;
    if ((rc = upload(ctx, &S->ip, ip)) || (rc = upload(ctx, &S->col, col)) || (rc = upload(ctx, &S->diag, dg)) ||
        (rc = upload(ctx, &S->a_in, a)) || (rc = upload(ctx, &S->lrows, lrows)) || (rc = upload(ctx, &S->loff, loff)) ||
        (rc = upload(ctx, &S->urows, urows)) || (rc = upload(ctx, &S->uoff, uoff)) ||
        (rc = upload(ctx, &S->f64, a)) || (rc = upload(ctx, &S->lpk_off, lpo)) || (rc = upload(ctx, &S->lpk_col, lpc)) ||
        (rc = upload(ctx, &S->lpk_src, lps)) || (rc = upload(ctx, &S->upk_off, upo)) ||
        (rc = upload(ctx, &S->upk_col, upc)) || (rc = upload(ctx, &S->upk_src, ups)) ||
        (rc = upload(ctx, &S->upk_dsrc, uds)) || (rc = workspace(S, 1))) {
        stokes_free(S);
        delete S;
        return rc;
    }
    S->l_nnz = l_nnz;
    S->u_nnz = u_nnz;
    const size_t ln = lpc.size(), un = upc.size();  // with the zero entries past the end
    if (hipMalloc((void**)&S->lpk64, ln * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&S->upk64, un * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&S->ud64, N * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&S->lpk32, ln * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&S->upk32, un * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&S->ud32, N * sizeof(float)) != hipSuccess) {
        stokes_free(S);
        delete S;
        return rbfx_fail(ctx, RBFX_E_NOMEM, "stokes_create: packed factor entries");
    }
    hipError_t e = hipMemsetAsync(S->lpk64, 0, ln * sizeof(double), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(S->upk64, 0, un * sizeof(double), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(S->lpk32, 0, ln * sizeof(float), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(S->upk32, 0, un * sizeof(float), ctx->stream);
    if (e == hipSuccess) {
        ilu0_factor_kernel<<<1, FACTOR_THREADS, 0, ctx->stream>>>(S->ip, S->col, S->diag, S->a_in, S->lrows, S->loff, nl,
                                                                  N * W, S->f64);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipStream_t st = ctx->stream;
        if (S->l_nnz) {
            gather_kernel<double><<<blocks_for(S->l_nnz, 256), 256, 0, st>>>(S->f64, S->lpk_src, S->lpk64, S->l_nnz);
            gather_kernel<float><<<blocks_for(S->l_nnz, 256), 256, 0, st>>>(S->f64, S->lpk_src, S->lpk32, S->l_nnz);
        }
        if (S->u_nnz) {
            gather_kernel<double><<<blocks_for(S->u_nnz, 256), 256, 0, st>>>(S->f64, S->upk_src, S->upk64, S->u_nnz);
            gather_kernel<float><<<blocks_for(S->u_nnz, 256), 256, 0, st>>>(S->f64, S->upk_src, S->upk32, S->u_nnz);
        }
        gather_kernel<double><<<blocks_for(N, 256), 256, 0, st>>>(S->f64, S->upk_dsrc, S->ud64, N);
        gather_kernel<float><<<blocks_for(N, 256), 256, 0, st>>>(S->f64, S->upk_dsrc, S->ud32, N);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        stokes_free(S);
        delete S;
        return rbfx_fail(ctx, RBFX_E_HIP, "stokes_create: ILU(0) factorization: %s", hipGetErrorString(e));
    }
    *out = S;
    return RBFX_OK;
}

void rbfx_stokes_destroy(rbfx_stokes* S) {
    if (!S) return;
    hipSetDevice(S->ctx->device);
    hipStreamSynchronize(S->ctx->stream);
    stokes_free(S);
    delete S;
}

int rbfx_stokes_info(const rbfx_stokes* S, int32_t* n_unknowns, int32_t* lower_levels, int32_t* upper_levels) {
    if (!S) return RBFX_E_INVALID;
    if (n_unknowns) *n_unknowns = (int32_t)S->n_tot;
    if (lower_levels) *lower_levels = S->n_lev_l;
    if (upper_levels) *upper_levels = S->n_lev_u;
    return RBFX_OK;
}

int rbfx_stokes_apply(rbfx_stokes* S, const double* d_x, double* d_y) {
    if (!S || !d_x || !d_y) return RBFX_E_INVALID;
    return apply<double>(S, d_x, d_y);
}

int rbfx_stokes_factor(rbfx_stokes* S, double* d_f) {
    if (!S || !d_f) return RBFX_E_INVALID;
    RBFX_HIP_CHECK(S->ctx, hipMemcpyAsync(d_f, S->f64, (size_t)S->N * S->width * sizeof(double), hipMemcpyDeviceToDevice,
                                          S->ctx->stream));
    return RBFX_OK;
}

int rbfx_stokes_precond(rbfx_stokes* S, double* d_v) {
    if (!S || !d_v) return RBFX_E_INVALID;
    return precond<double>(S, d_v);
}

int rbfx_stokes_gmres(rbfx_stokes* S, const double* d_b, double* d_x, double tol, int32_t max_iter, int32_t krylov,
                      int32_t mode, int32_t* iters, double* error) {
    if (!S || !d_b || !d_x || !iters || !error || !(tol > 0) || max_iter <= 0 || krylov <= 0 || krylov > max_iter)
        return RBFX_E_INVALID;
    rbfx_ctx* ctx = S->ctx;
    if (mode != RBFX_F64 && mode != RBFX_F32) return rbfx_fail(ctx, RBFX_E_INVALID, "stokes_gmres: mode %d", mode);
    if (mode == RBFX_F32 && !S->e32)
        return rbfx_fail(ctx, RBFX_E_STATE, "stokes_gmres: FP32 refinement needs the FP32 operator set");
    RBFX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    int rc = workspace(S, krylov);
    if (rc) return rc;
    *iters = 0;
    *error = 0.0;
    if (mode == RBFX_F64) {
        rc = gmres<double>(S, d_b, d_x, tol, max_iter, krylov, true, false, iters, error);
        return rc < 0 ? rc : RBFX_OK;
    }
    // FP32 inner solves (operator, factor and Krylov basis in FP32) inside FP64 refinement:
    // r = b - A x and the test ||M^-1 r|| / ||b|| < tol in FP64, the correction from one
    // FP32 GMRES(krylov) cycle that reduces its preconditioned residual by 1e-4, x += d.
    hipStream_t st = ctx->stream;
    const size_t n = S->n_tot;
    const int nb = blocks_for(n, 256);
    RBFX_HIP_CHECK(ctx, hipMemsetAsync(d_x, 0, n * sizeof(double), st));
    if ((rc = mdot<double>(S, d_b, 1, d_b, S->hbuf, 0))) return rc;
    double nb2 = 0.0;
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(&nb2, S->hbuf, sizeof(double), hipMemcpyDeviceToHost, st));
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(st));
    const double norm_b = std::sqrt(nb2);
    if (norm_b == 0.0) return RBFX_OK;
    float* r32 = (float*)S->r64;   // the FP32 vectors live in the FP64 buffers' space
    float* d32 = (float*)S->x32d;
    while (true) {
        if ((rc = apply<double>(S, d_x, (double*)S->w))) return rc;
        axpby_kernel<double, double><<<nb, 256, 0, st>>>((double*)S->w, d_b, 1.0, -1.0, n);  // w = b - A x
        RBFX_LAUNCH_CHECK(ctx, "axpby_kernel");
        to_f32_kernel<<<nb, 256, 0, st>>>((const double*)S->w, r32, n);
        RBFX_LAUNCH_CHECK(ctx, "to_f32_kernel");
        if ((rc = precond<double>(S, (double*)S->w))) return rc;
        if ((rc = mdot<double>(S, (double*)S->w, 1, (double*)S->w, S->hbuf, 0))) return rc;
        double p2 = 0.0;
        RBFX_HIP_CHECK(ctx, hipMemcpyAsync(&p2, S->hbuf, sizeof(double), hipMemcpyDeviceToHost, st));
        RBFX_HIP_CHECK(ctx, hipStreamSynchronize(st));
        *error = std::sqrt(p2) / norm_b;
        if (!(*error >= tol) || !std::isfinite(*error) || *iters >= max_iter) return RBFX_OK;
        const int budget = std::min<int>(krylov, max_iter - *iters);
        double inner_err = 0.0;
        const int before = *iters;
        rc = gmres<float>(S, r32, d32, 1e-4, budget, budget, true, true, iters, &inner_err);
        if (rc < 0) return rc;
        if (*iters == before) return RBFX_OK;  // the FP32 residual vanished: nothing left to correct
        axpby_kernel<double, float><<<nb, 256, 0, st>>>(d_x, d32, 1.0, 1.0, n);
        RBFX_LAUNCH_CHECK(ctx, "axpby_kernel");
    }
}

}  // extern "C"
