// Batched RBF-FD weight computation: one small dense saddle-point solve per stencil.
//
// Reference: RBFFD::computeAllWeightsForStencil_Direct (src/rbffd/rbffd.cpp:225-352) builds
//   M = [A P; P' 0], A(j,i) = phi_eps(x_i - x_j)   (distanceMatrix, rbffd.cpp:888-969)
//   R(:,t) = L_t phi(x_centre - x_j), L_t p(x_centre) for the polynomial rows (356-537)
// and calls arma::solve (LAPACK dgesv: LU with partial pivoting). HV columns are scaled by
// getHVScalar (295-297). Gaussian phi and its derivatives follow rbf_gaussian.h:40-249.
//
// GPU design. Stencils are grouped by system size so each launch sizes its layout for its
// class. Both solvers factor [M | R] by right-looking elimination with implicit partial
// pivoting (pivot = largest |a_rk| among the rows not yet used, found by a wave max over
// packed keys; the row order is kept in a permutation instead of moving rows), then
// back-substitute.
//  * FP64 (default): the register-resident solver, 1 to 8 waves per stencil with the matrix
//    in VGPRs (see "register-resident elimination" below).
//  * The one-wave LDS solver: [M | R] assembled straight into LDS, one row per lane, no block
//    barriers. It runs the FP32 factorisation of the refinement mode (x refined with FP64
//    residuals recomputed from the node coordinates; stencils whose refinement does not
//    converge are re-solved in FP64) and, with RBFX_WEIGHTS_F64=lds, the FP64 mode.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <mutex>
#include <type_traits>
#include <vector>

#include "rbfx_internal.h"

namespace {


constexpr int MAX_M = 128;
constexpr int MAX_T = 16;
constexpr size_t STATIC_LDS_RESERVE = 1024;  // static __shared__ scratch of the kernels (< 1 KB)

struct Mono3 {
    int8_t a, b, c;
};

// Monomials of degree 0..3 in centred coordinates; order: by degree, then descending power of
// x, then of y (the same list as the CPU restatement used by the tests).
__constant__ Mono3 c_mono[20] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {1, 1, 0}, {1, 0, 1},
                                 {0, 2, 0}, {0, 1, 1}, {0, 0, 2}, {3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0},
                                 {1, 1, 1}, {1, 0, 2}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}};

// Sphere-reduced basis (poly_deg | RBFX_POLY_SPHERE): on the unit sphere |d|^2 + 2 x0.d = 0,
// so the monomials with a power of z of 2 or more are dropped: 9 functions at degree 2, 16 at
// degree 3 (the dimensions of the spherical harmonics up to that degree).
__constant__ Mono3 c_mono_sph[16] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {1, 1, 0},
                                     {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {3, 0, 0}, {2, 1, 0}, {2, 0, 1},
                                     {1, 2, 0}, {1, 1, 1}, {0, 3, 0}, {0, 2, 1}};

int num_monomials(int poly_deg) {
    if (poly_deg < 0) return 0;
    int d = poly_deg & 15;
    return (poly_deg & RBFX_POLY_SPHERE) ? (d + 1) * (d + 1) : (d + 1) * (d + 2) * (d + 3) / 6;
}

// Degree used for a stencil of n nodes: the largest d <= poly_deg with at most n monomials
// (more monomials than nodes make [A P; P' 0] singular; degree 3 needs 20 nodes).
int effective_degree(int poly_deg, int n) {
    if (poly_deg < 0) return poly_deg;
    int flag = poly_deg & RBFX_POLY_SPHERE, d = poly_deg & 15;
    while (d > 0 && num_monomials(d | flag) > n) d--;
    return d | flag;
}

struct WParams {
    const int32_t* st_off;
    const int32_t* st_idx;
    const double* xyz;
    const double* eps;
    const int32_t* order;  // stencil ids of this size class
    int count;
    int m;                 // system size of the class
    int np;
    uint32_t types[MAX_T];
    int T;
    int dim, hv_k;
    double hv_scale;
    int nnz;
    double* w;
    double* diag;
    int* fail_flags;       // per stencil: 1 = singular / not converged (FP32 mode)
    int max_iter;
    int fp32;
    int sph;               // sphere-reduced polynomial basis
};

__device__ __forceinline__ double sq3(double x, double y, double z) { return x * x + y * y + z * z; }

// Gaussian RBF (rbf_gaussian.h) at diff = (x,y,z); e2 = eps^2.
__device__ double g_eval(double e2, double x, double y, double z) { return exp(-(sq3(x, y, z) * e2)); }

__device__ double g_hv_pk(double e2, int dim, int k, double r2) {
    // p_0 = 1, p_1 = 4 e2 r2 - 2d, p_k = 4(e2 r2 - 2(k-1) - d/2) p_{k-1} - 8(k-1)(2(k-1) - 2 + d) p_{k-2}
    double e2r2 = e2 * r2;
    if (k == 0) return 1.;
    double pm2 = 1., pm1 = 4. * (e2r2)-2. * dim;
    for (int q = 2; q <= k; q++) {
        double pk = 4. * (e2r2 - 2. * (q - 1) - dim / 2.) * pm1 - 8. * (q - 1) * (2. * (q - 1) - 2. + dim) * pm2;
        pm2 = pm1;
        pm1 = pk;
    }
    return pm1;
}

__device__ void g_sph(double x, double y, double z, double& lam, double& th) {
    lam = atan2(y, x);
    th = atan2(z, sqrt(x * x + y * y));
}

// R entry of stencil node j for operator `type` (rbffd.cpp:356-537), diff = x0 - xj.
__device__ double rbf_rhs(uint32_t type, double eps, double e2, int dim, int hv_k, const double* x0, const double* xj,
                          double lam0, double th0, double lamj, double thj) {
    double dx = x0[0] - xj[0], dy = x0[1] - xj[1], dz = x0[2] - xj[2];
    double ev = g_eval(e2, dx, dy, dz);
    switch (type) {
        case RBFX_X: return -2. * (dx * e2) * ev;
        case RBFX_Y: return -2. * (dy * e2) * ev;
        case RBFX_Z: return -2. * (dz * e2) * ev;
        case RBFX_LAPL:
            if (dim == 1) return 2. * e2 * (-1. + 2. * (dx * dx * e2)) * ev;
            if (dim == 2) return 4. * e2 * (-1. + dx * dx * e2 + dy * dy * e2) * ev;
            return (-6. * e2 + 4. * (sq3(dx, dy, dz) * (e2 * e2))) * ev;
        case RBFX_INTERP: return ev;
        case RBFX_R: return -2. * (sqrt(sq3(dx, dy, dz)) * e2) * ev;
        case RBFX_LAMBDA: {
            double d = cos(thj) * sin(lam0 - lamj);
            return (fabs(d) > 0.) ? d * (-2. * e2 * ev) : 0.;
        }
        case RBFX_THETA: {
            double d = cos(thj) * sin(th0) * cos(lam0 - lamj) - sin(thj) * cos(th0);
            return (fabs(d) > 0.) ? d * (-2. * e2 * ev) : 0.;
        }
        case RBFX_LSFC: {
            double r2 = sq3(dx, dy, dz);
            double dr = -2. * e2 * ev;
            double d2 = dr + 4. * (e2 * e2) * r2 * exp(-(e2 * r2));
            return (1. / 4.) * ((4. - r2) * d2 + (4. - 3. * r2) * dr);
        }
        case RBFX_XSFC: case RBFX_YSFC: case RBFX_ZSFC: {
            double xtx = x0[0] * xj[0] + x0[1] * xj[1] + x0[2] * xj[2];
            int c = (type == RBFX_XSFC) ? 0 : (type == RBFX_YSFC ? 1 : 2);
            return (x0[c] * xtx - xj[c]) * (-2. * e2 * ev);
        }
        case RBFX_HV: return pow(eps, 2 * hv_k) * g_hv_pk(e2, dim, hv_k, sq3(dx, dy, dz)) * ev;
        default: return 0.;
    }
}

__device__ double mono_eval(int p, int sph, double dx, double dy, double dz) {
    Mono3 mm = sph ? c_mono_sph[p] : c_mono[p];
    double v = 1.0;
    for (int i = 0; i < mm.a; i++) v *= dx;
    for (int i = 0; i < mm.b; i++) v *= dy;
    for (int i = 0; i < mm.c; i++) v *= dz;
    return v;
}

// Operator applied to monomial p at the centre (d = 0): only degree 0..2 terms contribute.
__device__ double mono_rhs(uint32_t type, int p, int sph, const double* x0, int dim, int hv_k, double lam0, double th0) {
    Mono3 mm = sph ? c_mono_sph[p] : c_mono[p];
    int deg = mm.a + mm.b + mm.c;
    double p0 = (deg == 0) ? 1.0 : 0.0;
    double g[3] = {0, 0, 0};
    double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (deg == 1) {
        g[0] = mm.a; g[1] = mm.b; g[2] = mm.c;
    } else if (deg == 2) {
        int e[3] = {mm.a, mm.b, mm.c};
        for (int i = 0; i < 3; i++) {
            if (e[i] == 2) H[i][i] = 2.0;
            for (int j = i + 1; j < 3; j++)
                if (e[i] == 1 && e[j] == 1) H[i][j] = H[j][i] = 1.0;
        }
    }
    double tr = 0;
    for (int i = 0; i < dim && i < 3; i++) tr += H[i][i];
    switch (type) {
        case RBFX_X: return g[0];
        case RBFX_Y: return g[1];
        case RBFX_Z: return g[2];
        case RBFX_LAPL: return tr;
        case RBFX_HV: return (hv_k == 0) ? p0 : (hv_k == 1 ? tr : 0.0);
        case RBFX_LAMBDA: return -sin(lam0) * g[0] + cos(lam0) * g[1];
        case RBFX_THETA: return -sin(th0) * cos(lam0) * g[0] - sin(th0) * sin(lam0) * g[1] + cos(th0) * g[2];
        case RBFX_LSFC: {
            double full = H[0][0] + H[1][1] + H[2][2];
            double xhx = 0, xg = 0;
            for (int i = 0; i < 3; i++) {
                xg += x0[i] * g[i];
                for (int j = 0; j < 3; j++) xhx += x0[i] * H[i][j] * x0[j];
            }
            return full - xhx - 2.0 * xg;
        }
        case RBFX_XSFC: case RBFX_YSFC: case RBFX_ZSFC: {
            int c = (type == RBFX_XSFC) ? 0 : (type == RBFX_YSFC ? 1 : 2);
            double xg = x0[0] * g[0] + x0[1] * g[1] + x0[2] * g[2];
            return g[c] - x0[c] * xg;
        }
        case RBFX_INTERP: return p0;
        default: return 0.0;
    }
}

// ---- one wave per system -----------------------------------------------------------------
// Each system is solved by one 64-lane workgroup with no block barriers: lane L owns rows L
// and L + 64 (m <= 128) of the augmented matrix [M | R] in LDS, so a rank-1 update is one
// multiply-add per owned row and column, the pivot row is read by every lane at the same
// address (an LDS broadcast), and all LDS traffic of a wave is ordered by the hardware.
// Several systems share a CU through the LDS they need (about 33 KB at m = 59 in FP64).
//
// LDS layout of one system:
// [X: n*3 f64][SPH: n*2 f64][XS: m*T f64][BS: m*T f64][perm: m i32][Aug: m*ld of TA]
// Aug = [M | R] with ld = m + T (+1 when even: rows start on alternating banks, so the 64
// lanes reading one column hit distinct banks). XS/BS (solution by unknown, original RHS by
// row) exist only in FP32 mode.
constexpr int WAVE = 64;
constexpr int UCH = 8;  // columns per chunk of the rank-1 update (loads batched ahead of stores)

__host__ __device__ inline int aug_ld(int m, int T) {
    int ld = m + T;
    return (ld & 1) ? ld : ld + 1;
}

__host__ __device__ inline size_t lds_head(int m, int n, int T, bool ir) {
    size_t b = (size_t)n * 5 * sizeof(double) + (ir ? 2 * (size_t)m * T * sizeof(double) : 0) + (size_t)m * sizeof(int);
    return (b + 15) / 16 * 16;
}

template <typename TA>
__host__ __device__ inline size_t lds_bytes(int m, int n, int T, bool ir) {
    return lds_head(m, n, T, ir) + (size_t)m * aug_ld(m, T) * sizeof(TA);
}

// Entry (r, c) of the FP64 saddle-point matrix, from the staged coordinates.
__device__ __forceinline__ double m_entry(int r, int c, int n, double e2, const double* X, int sph) {
    if (r < n && c < n)
        return g_eval(e2, X[3 * c] - X[3 * r], X[3 * c + 1] - X[3 * r + 1], X[3 * c + 2] - X[3 * r + 2]);
    if (r >= n && c >= n) return 0.0;
    int i = (r < n) ? r : c;
    int p = (r < n) ? c - n : r - n;
    return mono_eval(p, sph, X[3 * i] - X[0], X[3 * i + 1] - X[1], X[3 * i + 2] - X[2]);
}

struct Sys {
    const WParams* P;
    int s, off, n, m, T, ld;
    double eps, e2;
    bool sph;
    double* X;
    double* SPH;
    double* XS;
    double* BS;
    int* perm;
};

__device__ __forceinline__ double rhs_entry(const Sys& S, int r, int t) {
    const WParams& P = *S.P;
    double lam0 = S.sph ? S.SPH[0] : 0., th0 = S.sph ? S.SPH[1] : 0.;
    if (r < S.n)
        return rbf_rhs(P.types[t], S.eps, S.e2, P.dim, P.hv_k, S.X, S.X + 3 * r, lam0, th0, S.sph ? S.SPH[2 * r] : 0.,
                       S.sph ? S.SPH[2 * r + 1] : 0.);
    return mono_rhs(P.types[t], r - S.n, P.sph, S.X, P.dim, P.hv_k, lam0, th0);
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffull), off, WAVE);
        unsigned hi = __shfl_xor((unsigned)(v >> 32), off, WAVE);
        unsigned long long o = ((unsigned long long)hi << 32) | lo;
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
    return v;
}

// Pivot key of a candidate row: |a| with its lowest 8 mantissa bits replaced by 255 - r, so
// one unsigned 64-bit max picks the largest |a| (to within 2^-44 relative) and, among equal
// keys, the smaller row. Rows already used as pivots do not take part (key 0).
__device__ __forceinline__ unsigned long long pivot_key(double a, int r) {
    unsigned long long b = (unsigned long long)__double_as_longlong(fabs(a));
    return (b & ~0xffull) | (unsigned long long)(255 - r);
}

// Right-looking elimination of [M | R] with implicit partial pivoting. perm[k] is the pivot
// row of step k; posr[h] is the step at which this lane's row h was used as pivot (m = not
// yet). The multiplier of row r at step k is kept in Aug[r][k] so that later right-hand
// sides can be eliminated. Returns false on an exact zero pivot (uniform over the wave).
template <typename TA>
__device__ bool factor(const Sys& S, TA* Aug, int (&posr)[2]) {
    const int lane = threadIdx.x, m = S.m, ld = S.ld, wcols = m + S.T;
    const int r0 = lane, r1 = lane + WAVE;
    posr[0] = posr[1] = m;
    for (int k = 0; k < m; k++) {
        unsigned long long key = 0;
        if (r0 < m && posr[0] == m) key = pivot_key((double)Aug[r0 * ld + k], r0);
        if (r1 < m && posr[1] == m) {
            unsigned long long k1 = pivot_key((double)Aug[r1 * ld + k], r1);
            key = k1 > key ? k1 : key;
        }
        key = wave_max_u64(key);
        const int p = 255 - (int)(key & 0xffull);
        const TA piv = Aug[p * ld + k];
        if (piv == (TA)0) return false;
        if (lane == 0) S.perm[k] = p;
        if (r0 == p) posr[0] = k;
        if (r1 == p) posr[1] = k;
        const bool a0 = r0 < m && posr[0] == m, a1 = r1 < m && posr[1] == m;
        const TA inv = (TA)1 / piv;
        TA l0 = 0, l1 = 0;
        if (a0) { l0 = Aug[r0 * ld + k] * inv; Aug[r0 * ld + k] = l0; }
        if (a1) { l1 = Aug[r1 * ld + k] * inv; Aug[r1 * ld + k] = l1; }
        const TA* prow = Aug + p * ld;
        TA* row0 = Aug + r0 * ld;
        TA* row1 = Aug + r1 * ld;
        int c = k + 1;
        for (; c + UCH <= wcols; c += UCH) {
            TA pv[UCH], v0[UCH], v1[UCH];
#pragma unroll
            for (int q = 0; q < UCH; q++) pv[q] = prow[c + q];
            if (a0) {
#pragma unroll
                for (int q = 0; q < UCH; q++) v0[q] = row0[c + q];
            }
            if (a1) {
#pragma unroll
                for (int q = 0; q < UCH; q++) v1[q] = row1[c + q];
            }
            if (a0) {
#pragma unroll
                for (int q = 0; q < UCH; q++) row0[c + q] = v0[q] - l0 * pv[q];
            }
            if (a1) {
#pragma unroll
                for (int q = 0; q < UCH; q++) row1[c + q] = v1[q] - l1 * pv[q];
            }
        }
        for (; c < wcols; c++) {
            TA pv = prow[c];
            if (a0) row0[c] = row0[c] - l0 * pv;
            if (a1) row1[c] = row1[c] - l1 * pv;
        }
    }
    return true;
}

// Back substitution on the eliminated RHS columns: the solution of unknown k overwrites
// Aug[perm[k]][m + t]. Rows pivoted before step k subtract their U(., k) x_k.
template <typename TA>
__device__ void back_solve(const Sys& S, TA* Aug, const int (&posr)[2]) {
    const int lane = threadIdx.x, m = S.m, ld = S.ld, T = S.T;
    const int r0 = lane, r1 = lane + WAVE;
    for (int k = m - 1; k >= 0; k--) {
        const int p = S.perm[k];
        const TA ukk = Aug[p * ld + k];
        for (int t = 0; t < T; t++) {
            TA x = Aug[p * ld + m + t] / ukk;
            if (r0 < m && posr[0] < k) Aug[r0 * ld + m + t] -= Aug[r0 * ld + k] * x;
            if (r1 < m && posr[1] < k) Aug[r1 * ld + m + t] -= Aug[r1 * ld + k] * x;
            if ((r0 == p) | (r1 == p)) Aug[p * ld + m + t] = x;
        }
    }
}

// Forward elimination of new RHS columns (rows in original order) with the stored multipliers.
template <typename TA>
__device__ void forward_solve(const Sys& S, TA* Aug, const int (&posr)[2]) {
    const int lane = threadIdx.x, m = S.m, ld = S.ld, T = S.T;
    const int r0 = lane, r1 = lane + WAVE;
    for (int k = 0; k < m - 1; k++) {
        const int p = S.perm[k];
        for (int t = 0; t < T; t++) {
            TA y = Aug[p * ld + m + t];
            if (r0 < m && posr[0] > k) Aug[r0 * ld + m + t] -= Aug[r0 * ld + k] * y;
            if (r1 < m && posr[1] > k) Aug[r1 * ld + m + t] -= Aug[r1 * ld + k] * y;
        }
    }
}

// r = b - M x in FP64 with M rebuilt from the coordinates; x(k,t) = XS[k*T+t] (by unknown).
// With Aug set, r is written into its RHS columns (as TA). Returns the normwise backward
// error max_t ||r_t|| / (||M|| ||x_t|| + ||b_t||) (max norms, ||M|| = max row sum). The
// right-hand sides are taken RC at a time (M is rebuilt once per group) so that the
// accumulators stay in registers with compile-time indices; the operations of each r_t are
// the same in any grouping.
// TAG separates instances by caller (each instance gets its caller's register budget).
template <typename TA, int MT, int TAG = 0>
__device__ double residual(const Sys& S, TA* Aug, const double* XS) {
    constexpr int RC = 4;
    const int lane = threadIdx.x & (WAVE - 1), m = S.m, ld = S.ld, T = S.T;
    double be = 0.0;
    for (int t0 = 0; t0 < T; t0 += RC) {
        double rmax[RC], xmax[RC], bmax[RC];
#pragma unroll
        for (int t = 0; t < RC; t++) rmax[t] = xmax[t] = bmax[t] = 0.0;
        double rowsum = 0.0;
        for (int h = 0; h < 2; h++) {
            const int r = lane + h * WAVE;
            if (r >= m) continue;
            double acc[RC];
#pragma unroll
            for (int t = 0; t < RC; t++) {
                double b = 0.0;
                if (t0 + t < T) b = S.BS ? S.BS[r * T + t0 + t] : rhs_entry(S, r, t0 + t);
                acc[t] = b;
                bmax[t] = fmax(bmax[t], fabs(b));
            }
            double rs = 0.0;
            for (int c = 0; c < m; c++) {
                double a = m_entry(r, c, S.n, S.e2, S.X, S.P->sph);
                rs += fabs(a);
#pragma unroll
                for (int t = 0; t < RC; t++)
                    if (t0 + t < T) acc[t] -= a * XS[c * T + t0 + t];
            }
            rowsum = fmax(rowsum, rs);
#pragma unroll
            for (int t = 0; t < RC; t++) {
                if (t0 + t >= T) continue;
                rmax[t] = fmax(rmax[t], fabs(acc[t]));
                xmax[t] = fmax(xmax[t], fabs(XS[r * T + t0 + t]));
                if (Aug) Aug[r * ld + m + t0 + t] = (TA)acc[t];
            }
        }
        rowsum = wave_max(rowsum);
#pragma unroll
        for (int t = 0; t < RC; t++)
            if (t0 + t < T) be = fmax(be, wave_max(rmax[t]) / (rowsum * wave_max(xmax[t]) + wave_max(bmax[t])));
    }
    return be;
}

// max|U| over the factored rows, for the growth factor max|U| / max|M|.
template <typename TA>
__device__ double umax_of(const Sys& S, const TA* Aug, const int (&posr)[2]) {
    const int lane = threadIdx.x, m = S.m, ld = S.ld;
    double u = 0.0;
    for (int h = 0; h < 2; h++) {
        const int r = lane + h * WAVE;
        if (r >= m) continue;
        for (int c = posr[h]; c < m; c++) u = fmax(u, fabs((double)Aug[r * ld + c]));
    }
    return wave_max(u);
}

__device__ void stage_nodes(Sys& S) {
    const WParams& P = *S.P;
    for (int j = threadIdx.x; j < S.n; j += WAVE) {
        const double* p = P.xyz + 3 * (size_t)P.st_idx[S.off + j];
        S.X[3 * j] = p[0];
        S.X[3 * j + 1] = p[1];
        S.X[3 * j + 2] = p[2];
        if (S.sph) {
            double lam, th;
            g_sph(p[0], p[1], p[2], lam, th);
            S.SPH[2 * j] = lam;
            S.SPH[2 * j + 1] = th;
        }
    }
}

template <typename TA>
__device__ TA* setup(Sys& S, const WParams& P, unsigned char* smem, int s, bool ir) {
    S.P = &P;
    S.s = s;
    S.off = P.st_off[s];
    S.n = P.st_off[s + 1] - S.off;
    S.m = P.m;
    S.T = P.T;
    S.ld = aug_ld(S.m, S.T);
    S.eps = P.eps[s];
    S.e2 = S.eps * S.eps;
    S.sph = false;
    for (int t = 0; t < S.T; t++) S.sph |= (P.types[t] == RBFX_LAMBDA || P.types[t] == RBFX_THETA);
    double* d = (double*)smem;
    S.X = d;
    S.SPH = S.X + 3 * S.n;
    double* q = S.SPH + 2 * S.n;
    if (ir) {
        S.XS = q;
        S.BS = q + S.m * S.T;
        q += 2 * S.m * S.T;
    } else {
        S.XS = S.BS = nullptr;
    }
    S.perm = (int*)q;
    return (TA*)(smem + lds_head(S.m, S.n, S.T, ir));
}

// Assembles [M | R] (M in TA, R in TA and, in FP32 mode, FP64 in BS) row by row, one row per
// lane; returns max|M| over the wave.
template <typename TA>
__device__ double assemble(const Sys& S, TA* Aug) {
    const int lane = threadIdx.x, m = S.m, ld = S.ld, T = S.T;
    double amax = 0.0;
    for (int h = 0; h < 2; h++) {
        const int r = lane + h * WAVE;
        if (r >= m) continue;
        for (int c = 0; c < m; c++) {
            double a = m_entry(r, c, S.n, S.e2, S.X, S.P->sph);
            amax = fmax(amax, fabs(a));
            Aug[r * ld + c] = (TA)a;
        }
        for (int t = 0; t < T; t++) {
            double b = rhs_entry(S, r, t);
            Aug[r * ld + m + t] = (TA)b;
            if (S.BS) {
                S.BS[r * T + t] = b;
                S.XS[r * T + t] = 0.0;
            }
        }
    }
    return wave_max(amax);
}

__device__ void write_failed(const Sys& S, const WParams& P) {
    for (int q = threadIdx.x; q < S.n * S.T; q += WAVE) {
        int j = q / S.T, t = q - j * S.T;
        P.w[(size_t)t * P.nnz + S.off + j] = NAN;
    }
    if (threadIdx.x == 0) {
        P.fail_flags[S.s] = 1;
        if (P.diag) {
            P.diag[4 * (size_t)S.s] = NAN;
            P.diag[4 * (size_t)S.s + 1] = 0;
            P.diag[4 * (size_t)S.s + 2] = 0;
            P.diag[4 * (size_t)S.s + 3] = RBFX_PATH_SINGULAR;
        }
    }
}

// FP64 path: assemble, factor, back-substitute (the arma::solve restatement on the GPU).
// MT bounds the number of right-hand sides (register arrays of the diagnostics).
template <int MT>
__global__ void __launch_bounds__(WAVE) weights_f64_kernel(WParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    if ((int)blockIdx.x >= P.count) return;
    const int s = P.order[blockIdx.x];
    Sys S;
    double* Aug = setup<double>(S, P, smem, s, false);
    const int lane = threadIdx.x, m = S.m, T = S.T, ld = S.ld;
    stage_nodes(S);
    const double amax = assemble<double>(S, Aug);
    int posr[2];
    if (!factor<double>(S, Aug, posr)) {
        write_failed(S, P);
        return;
    }
    back_solve<double>(S, Aug, posr);
    for (int j = lane; j < S.n; j += WAVE) {
        const int pj = S.perm[j];
        for (int t = 0; t < T; t++) {
            double scale = (P.types[t] == RBFX_HV) ? P.hv_scale : 1.0;
            P.w[(size_t)t * P.nnz + S.off + j] = Aug[pj * ld + m + t] * scale;
        }
    }
    if (lane == 0) P.fail_flags[s] = 0;
    if (P.diag) {
        const double gr = umax_of<double>(S, Aug, posr) / amax;
        // x by unknown into the (no longer needed) coordinate-free head of Aug's row space:
        // read all values first, then write, so no lane overwrites a value another still reads
        double* xs = (double*)(smem + lds_head(m, S.n, T, false)) + (size_t)m * ld - (size_t)m * T;
        double tmp[2][MT];
        for (int h = 0; h < 2; h++) {
            const int k = lane + h * WAVE;
            if (k < m)
                for (int t = 0; t < T; t++) tmp[h][t] = Aug[S.perm[k] * ld + m + t];
        }
        __builtin_amdgcn_wave_barrier();
        for (int h = 0; h < 2; h++) {
            const int k = lane + h * WAVE;
            if (k < m)
                for (int t = 0; t < T; t++) xs[k * T + t] = tmp[h][t];
        }
        __builtin_amdgcn_wave_barrier();
        const double be = residual<double, MT>(S, (double*)nullptr, xs);
        if (lane == 0) {
            P.diag[4 * (size_t)s] = be;
            P.diag[4 * (size_t)s + 1] = gr;
            P.diag[4 * (size_t)s + 2] = 0;
            P.diag[4 * (size_t)s + 3] = RBFX_PATH_F64;
        }
    }
}

// ---- register-resident elimination (FP64) -------------------------------------------------
// The augmented matrix [M | R] lives in registers: lane L of each of the W waves of a block
// holds rows L and L + 64 (R = 2; R = 1 when m <= 64) of the columns c = W*j + w that wave w
// owns (slot j = 0..NJ-1, column-cyclic; slots past m + T hold zeros), so a rank-1 update is
// R FMAs per owned column and lane, with no LDS traffic for the matrix. Elimination step k:
//   * column k, stored to LDS by its owner wave at the end of step k-1, is read by every
//     wave, which finds the pivot itself (a DPP max over packed keys: same data, same result);
//   * the pivot lane of each wave copies its row's slots to a per-wave LDS row buffer, which
//     the wave reads back as 16-byte broadcasts;
//   * the update runs over groups of 8 slots, skipping the groups that hold only columns
//     <= k. Columns <= k of unpivoted rows are dead (the pivot search and the multipliers only
//     read column k), so the slots of a live group are updated without a per-column test;
//     pivoted rows have multiplier 0 and keep their values;
//   * the owner wave of column k+1 stores the update group that holds it (slots with
//     compile-time indices; the readers pick the column's slot at run time), double-buffered;
//     one block barrier ends the step.
// After the last step each lane stores its row's U part (columns > its pivot step) and its
// eliminated right-hand sides to LDS; back substitution then runs in pivot order, each wave on
// its own right-hand sides with lane i holding unknown i. Every live element sees the
// operations of the one-wave LDS kernel in the same order, so the two give bit-identical
// weights.
constexpr int REG_G = 8;  // slots per update group

// LDS: [pivot rows: W*NJ][column groups: 2 x rows x REG_G][X: 3n][SPH: 2n][pivots: m]
// [U: m(m+1)/2][BS: m*T] doubles (rows = 64 R); column c of U (rows 0..c in pivot order)
// starts at c(c+1)/2.
__host__ __device__ inline size_t reg_lds_bytes(int m, int n, int T, int wnj, int rows) {
    return ((size_t)wnj + 2 * (size_t)rows * REG_G + (size_t)n * 5 + m + (size_t)m * (m + 1) / 2 + (size_t)m * T) *
           sizeof(double);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ unsigned umax32_dpp_step(unsigned v) {
    const unsigned o = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROWMASK, 0xf, false);
    return o > v ? o : v;
}

// Maximum of a 32-bit value over the 64 lanes as a wave-uniform value (all lanes must be
// active): two quad swaps and two row rotations give every lane its 16-lane row's maximum, two
// row broadcasts fold rows 0..2 into row 3, and lane 63 is read.
__device__ __forceinline__ unsigned wave_umax32_uniform(unsigned v) {
    v = umax32_dpp_step<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
    v = umax32_dpp_step<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
    v = umax32_dpp_step<0x124, 0xf>(v);  // row_ror:4
    v = umax32_dpp_step<0x128, 0xf>(v);  // row_ror:8
    v = umax32_dpp_step<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = umax32_dpp_step<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Maximum of a 64-bit key over the 64 lanes, wave-uniform: the largest high word, then the
// largest low word among the lanes holding it (the same order as a 64-bit compare).
__device__ __forceinline__ unsigned long long wave_umax_uniform(unsigned long long v) {
    const unsigned hi = (unsigned)(v >> 32), lo = (unsigned)v;
    const unsigned mh = wave_umax32_uniform(hi);
    const unsigned ml = wave_umax32_uniform(hi == mh ? lo : 0u);
    return ((unsigned long long)mh << 32) | ml;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Makes this wave's earlier LDS stores visible to its later LDS loads in other lanes: without
// it the compiler may run the loads of lanes that did not store ahead of the stores of the one
// that did (lanes of a wave do not synchronise in the memory model).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Elimination step with the pivot row in slot H of lane pl; the groups below G0 hold only
// columns <= k (a compile-time bound, so every update is in place). The owner wave of column
// k+1 (own) stores group G0, which holds that column, row by row into cg.
template <int R, int NJ, int H, int G0>
__device__ __forceinline__ void reg_step(double (&a)[R][NJ], const double (&l)[R], bool own, int lane, int pl,
                                         double2* rb, double* cg) {
    constexpr int NG = NJ / REG_G, GQ = REG_G / 2;
    wave_lds_sync();  // the previous step's row reads are done before the row is replaced
    if (lane == pl) {
#pragma unroll
        for (int q = G0 * GQ; q < NJ / 2; q++) rb[q] = make_double2(a[H][2 * q], a[H][2 * q + 1]);
    }
    wave_lds_sync();
    // the row values of group g+1 are loaded while group g is updated; the scheduling
    // barriers keep the loads of later groups from being hoisted (one group of lookahead)
    double2 un[GQ];
#pragma unroll
    for (int q = 0; q < GQ; q++) un[q] = rb[G0 * GQ + q];
#pragma unroll
    for (int g = G0; g < NG; g++) {
        double2 u[GQ];
#pragma unroll
        for (int q = 0; q < GQ; q++) u[q] = un[q];
        if (g + 1 < NG) {
#pragma unroll
            for (int q = 0; q < GQ; q++) un[q] = rb[(g + 1) * GQ + q];
        }
#pragma unroll
        for (int q = 0; q < GQ; q++) {
            const int j = g * REG_G + 2 * q;
#pragma unroll
            for (int h = 0; h < R; h++) {
                a[h][j] = a[h][j] - l[h] * u[q].x;
                a[h][j + 1] = a[h][j + 1] - l[h] * u[q].y;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if (own) {
#pragma unroll
        for (int h = 0; h < R; h++) {
            double2* dst = (double2*)(cg + (lane + WAVE * h) * REG_G);
#pragma unroll
            for (int q = 0; q < GQ; q++) dst[q] = make_double2(a[h][G0 * REG_G + 2 * q], a[h][G0 * REG_G + 2 * q + 1]);
        }
    }
}

// Elimination steps of phase G: the steps k with W*REG_G*G <= k + 1 - w < W*REG_G*(G+1), in
// which the groups below G hold only dead columns (phases differ between waves; every step
// ends with one block barrier in every wave). Returns false on an exact zero pivot.
template <int W, int R, int NJ, int G>
__device__ __forceinline__ bool reg_phase(double (&a)[R][NJ], int (&posr)[R], int& k, int m, int wv, int lane, int tid, double* col,
                          double* dg, double2* rb) {
    const int kend = (G + 1 < NJ / REG_G) ? min(m, W * REG_G * (G + 1) + wv - 1) : m;
    for (; k < kend; k++) {
        const double* ck = col + (k & 1) * (R * WAVE * REG_G) + (k / W) % REG_G;  // column k, row 0
        double cv[R];
        unsigned long long key = 0;
#pragma unroll
        for (int h = 0; h < R; h++) {
            const int r = lane + WAVE * h;
            cv[h] = (r < m) ? ck[r * REG_G] : 0.0;
            if (r < m && posr[h] == m) {
                const unsigned long long kh = pivot_key(cv[h], r);
                key = kh > key ? kh : key;
            }
        }
        key = wave_umax_uniform(key);
        const int p = 255 - (int)(key & 0xffull);
        const int pl = p & (WAVE - 1), ph = p / WAVE;
        const double piv = readlane_f64((R == 1 || ph == 0) ? cv[0] : cv[R - 1], pl);
        if (piv == 0.0) return false;  // the same LDS data in every wave: a uniform exit
#pragma unroll
        for (int h = 0; h < R; h++)
            if (lane + WAVE * h == p) posr[h] = k;
        if (tid == 0) dg[k] = piv;
        const double inv = 1.0 / piv;
        double l[R];
#pragma unroll
        for (int h = 0; h < R; h++) l[h] = (lane + WAVE * h < m && posr[h] == m) ? cv[h] * inv : 0.0;
        const bool own = (k + 1) % W == wv;  // column k+1 lies in group G of this wave
        double* cg = col + ((k + 1) & 1) * (R * WAVE * REG_G);
        if (R == 1 || ph == 0)
            reg_step<R, NJ, 0, G>(a, l, own, lane, pl, rb, cg);
        else
            reg_step<R, NJ, R - 1, G>(a, l, own, lane, pl, rb, cg);
        if (W > 1) __syncthreads();
    }
    if constexpr (G + 1 < NJ / REG_G)
        return reg_phase<W, R, NJ, G + 1>(a, posr, k, m, wv, lane, tid, col, dg, rb);
    else
        return true;
}

// DIAG: the variant that also writes the backward error and growth factor (its residual
// code needs registers the plain variant does without).
template <int W, int R, int NJ, bool DIAG>
__global__ void __launch_bounds__(WAVE * W) weights_f64_reg_kernel(WParams P) {
    static_assert(NJ % REG_G == 0, "slots come in whole update groups");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double s_amax[W], s_umax[W];
    if ((int)blockIdx.x >= P.count) return;
    const int s = P.order[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane(tid / WAVE);  // wave-uniform column owner
    Sys S;
    S.P = &P;
    S.s = s;
    S.off = P.st_off[s];
    S.n = P.st_off[s + 1] - S.off;
    S.m = P.m;
    S.T = P.T;
    S.ld = 0;
    S.eps = P.eps[s];
    S.e2 = S.eps * S.eps;
    S.sph = false;
    for (int t = 0; t < S.T; t++) S.sph |= (P.types[t] == RBFX_LAMBDA || P.types[t] == RBFX_THETA);
    const int m = S.m, T = S.T, wcols = m + T;
    double2* rb = (double2*)smem + wv * (NJ / 2);  // this wave's pivot-row buffer
    double* col = (double*)smem + W * NJ;          // [2][64 R][REG_G]: the group holding column k
    S.X = col + 2 * R * WAVE * REG_G;
    S.SPH = S.X + 3 * S.n;
    S.XS = S.BS = nullptr;
    S.perm = nullptr;
    double* dg = S.SPH + 2 * S.n;  // pivots u_kk
    double* U = dg + m;            // packed upper triangle by columns
    double* BS = U + (size_t)m * (m + 1) / 2;  // [m][T]: staged R by row, later R' by pivot order
    for (int j = tid; j < S.n; j += WAVE * W) {
        const double* q = P.xyz + 3 * (size_t)P.st_idx[S.off + j];
        S.X[3 * j] = q[0];
        S.X[3 * j + 1] = q[1];
        S.X[3 * j + 2] = q[2];
        if (S.sph) {
            double lam, th;
            g_sph(q[0], q[1], q[2], lam, th);
            S.SPH[2 * j] = lam;
            S.SPH[2 * j + 1] = th;
        }
    }
    __syncthreads();
    for (int q = tid; q < m * T; q += WAVE * W) BS[q] = rhs_entry(S, q / T, q - (q / T) * T);
    __syncthreads();
    double a[R][NJ];
    double amax = 0.0;
#pragma unroll
    for (int h = 0; h < R; h++) {
        const int r = lane + WAVE * h;
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int c = W * j + wv;
            double v = 0.0;
            if (c < m) {
                if (r < m) {
                    v = m_entry(r, c, S.n, S.e2, S.X, P.sph);
                    amax = fmax(amax, fabs(v));
                }
            } else if (c < wcols) {
                if (r < m) v = BS[r * T + (c - m)];
            }
            a[h][j] = v;
            __builtin_amdgcn_sched_barrier(0);  // one entry at a time: short live ranges
        }
    }
    amax = wave_max(amax);
    if (lane == 0) s_amax[wv] = amax;
    int posr[R];
#pragma unroll
    for (int h = 0; h < R; h++) {
        posr[h] = m;
        if (wv == 0) {  // group 0 of wave 0 holds column 0
            double2* dst = (double2*)(col + (lane + WAVE * h) * REG_G);
#pragma unroll
            for (int q = 0; q < REG_G / 2; q++) dst[q] = make_double2(a[h][2 * q], a[h][2 * q + 1]);
        }
    }
    __syncthreads();  // column 0 stored; the staged R has been read
    int k = 0;
    const bool ok = reg_phase<W, R, NJ, 0>(a, posr, k, m, wv, lane, tid, col, dg, rb);
    if (!ok) {
        if (wv == 0) write_failed(S, P);
        return;
    }
    // U part (columns posr+1 .. m-1) and eliminated right-hand sides of each row, by pivot order
    double um = 0.0;
#pragma unroll
    for (int h = 0; h < R; h++) {
        const int i = posr[h];
        if (lane + WAVE * h >= m) continue;
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int c = W * j + wv;
            if (c < m) {
                if (c >= i) um = fmax(um, fabs(a[h][j]));
                if (c > i) U[(size_t)c * (c + 1) / 2 + i] = a[h][j];
            } else if (c < wcols) {
                BS[i * T + (c - m)] = a[h][j];
            }
        }
    }
    if (DIAG) {
        um = wave_max(um);
        if (lane == 0) s_umax[wv] = um;
    }
    __syncthreads();  // U, BS and the pivots complete
    // back substitution, right-hand sides t = wv + W*q; lane i holds unknown i (pivot order)
    constexpr int QT = (MAX_T + W - 1) / W;
    double b[QT][R];
#pragma unroll
    for (int h = 0; h < R; h++) {
        const int i = lane + WAVE * h;
#pragma unroll
        for (int q = 0; q < QT; q++) {
            const int t = wv + W * q;
            b[q][h] = (t < T && i < m) ? BS[i * T + t] : 0.0;
        }
    }
    for (int k = m - 1; k >= 0; k--) {
        const double* uc = U + (size_t)k * (k + 1) / 2;
        const double ukk = dg[k];
        double ucol[R];
#pragma unroll
        for (int h = 0; h < R; h++) ucol[h] = (lane + WAVE * h < k) ? uc[lane + WAVE * h] : 0.0;
#pragma unroll
        for (int q = 0; q < QT; q++) {
            if (wv + W * q >= T) continue;
            const double bk = readlane_f64((R == 1 || k < WAVE) ? b[q][0] : b[q][R - 1], k & (WAVE - 1));
            const double x = bk / ukk;
#pragma unroll
            for (int h = 0; h < R; h++) {
                const int i = lane + WAVE * h;
                if (i < k)
                    b[q][h] -= ucol[h] * x;
                else if (i == k)
                    b[q][h] = x;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < QT; q++) {
        const int t = wv + W * q;
        if (t >= T) continue;
        const double scale = (P.types[t] == RBFX_HV) ? P.hv_scale : 1.0;
#pragma unroll
        for (int h = 0; h < R; h++) {
            const int i = lane + WAVE * h;
            if (i < S.n) P.w[(size_t)t * P.nnz + S.off + i] = b[q][h] * scale;
        }
    }
    if (tid == 0) P.fail_flags[s] = 0;
    if (!DIAG) return;
    // diagnostics: x by unknown into BS (each wave overwrites only its own columns)
#pragma unroll
    for (int q = 0; q < QT; q++) {
        const int t = wv + W * q;
        if (t >= T) continue;
#pragma unroll
        for (int h = 0; h < R; h++) {
            const int i = lane + WAVE * h;
            if (i < m) BS[i * T + t] = b[q][h];
        }
    }
    __syncthreads();
    if (wv != 0) return;
    double am = 0.0, umx = 0.0;
    for (int w = 0; w < W; w++) {
        am = fmax(am, s_amax[w]);
        umx = fmax(umx, s_umax[w]);
    }
    const double be = residual<double, MAX_T>(S, (double*)nullptr, BS);
    if (lane == 0) {
        P.diag[4 * (size_t)s] = be;
        P.diag[4 * (size_t)s + 1] = umx / am;
        P.diag[4 * (size_t)s + 2] = 0;
        P.diag[4 * (size_t)s + 3] = RBFX_PATH_F64;
    }
}

// FP32 path: FP32 factorisation of [M | R], then x += LU^-1 (b - M x) with the residual in
// FP64 (M rebuilt in FP64, b kept in FP64) until the correction stops shrinking or is below
// FP64 resolution. A stencil that does not converge is flagged (fail_flags = 2) for an FP64
// re-solve.
template <int MT>
__global__ void __launch_bounds__(WAVE) weights_f32ir_kernel(WParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    if ((int)blockIdx.x >= P.count) return;
    const int s = P.order[blockIdx.x];
    Sys S;
    float* Aug = setup<float>(S, P, smem, s, true);
    const int lane = threadIdx.x, m = S.m, T = S.T, ld = S.ld;
    stage_nodes(S);
    const double amax = assemble<float>(S, Aug);
    int posr[2];
    if (!factor<float>(S, Aug, posr)) {
        if (lane == 0) P.fail_flags[s] = 2;  // FP32 pivot underflow: let FP64 decide
        return;
    }
    back_solve<float>(S, Aug, posr);
    double prev = INFINITY, be = INFINITY;
    int it = 0;
    bool ok = false;
    for (;;) {
        // x += d, where d(k,t) = Aug[perm[k]][m+t]
        double dmax = 0.0, xmax = 0.0;
        for (int h = 0; h < 2; h++) {
            const int k = lane + h * WAVE;
            if (k >= m) continue;
            const int p = S.perm[k];
            for (int t = 0; t < T; t++) {
                double d = (double)Aug[p * ld + m + t];
                double x = S.XS[k * T + t] + d;
                S.XS[k * T + t] = x;
                dmax = fmax(dmax, fabs(d));
                xmax = fmax(xmax, fabs(x));
            }
        }
        const double rel = wave_max(dmax) / wave_max(xmax);
        be = residual<float, MT>(S, Aug, S.XS);  // also loads r into the RHS columns
        it++;
        if (!(rel < 0.5 * prev) && it > 1) break;  // stagnation or divergence
        if (rel <= 2.2e-16 || be <= 4.0 * 1.1e-16) { ok = true; break; }
        if (it >= P.max_iter) break;
        prev = rel;
        forward_solve<float>(S, Aug, posr);
        back_solve<float>(S, Aug, posr);
    }
    ok = ok || be <= 4.0 * 1.1e-16 * m;
    if (!ok) {
        if (lane == 0) P.fail_flags[s] = 2;
        return;
    }
    for (int j = lane; j < S.n; j += WAVE)
        for (int t = 0; t < T; t++) {
            double scale = (P.types[t] == RBFX_HV) ? P.hv_scale : 1.0;
            P.w[(size_t)t * P.nnz + S.off + j] = S.XS[j * T + t] * scale;
        }
    if (lane == 0) P.fail_flags[s] = 0;
    if (P.diag) {
        const double gr = umax_of<float>(S, Aug, posr) / amax;
        if (lane == 0) {
            P.diag[4 * (size_t)s] = be;
            P.diag[4 * (size_t)s + 1] = gr;
            P.diag[4 * (size_t)s + 2] = it;
            P.diag[4 * (size_t)s + 3] = RBFX_PATH_F32_IR;
        }
    }
}

template <int W, int R, int NJ>
hipError_t launch_reg(int count, int n, hipStream_t st, const WParams& Q) {
    if (Q.m > R * WAVE || Q.m + Q.T > W * NJ) return hipErrorInvalidValue;
    const size_t lds = reg_lds_bytes(Q.m, n, Q.T, W * NJ, R * WAVE);
    if (Q.diag)
        hipLaunchKernelGGL((weights_f64_reg_kernel<W, R, NJ, true>), dim3(count), dim3(WAVE * W), lds, st, Q);
    else
        hipLaunchKernelGGL((weights_f64_reg_kernel<W, R, NJ, false>), dim3(count), dim3(WAVE * W), lds, st, Q);
    return hipGetLastError();
}

// Register-resident FP64 solve: waves W, rows per lane R (R = 1 for m <= 64) and column
// slots NJ (W * NJ >= m + T) by system size. Measured at 10^6 stencils per class, degree 2,
// 4 right-hand sides: m = 27 (1,1,32) 4.5 ms; m = 41 (2,1,24) 7.8 ms; m = 60 (2,1,32) 15.6 ms;
// m = 85 (4,2,24) 57 ms; m = 111 (8,2,16) 129 ms; wider or narrower layouts of each class
// were slower (fewer waves per SIMD, or more barrier-bound steps).
hipError_t launch_f64_reg(int m, int T, int count, int n, hipStream_t st, const WParams& Q) {
    const int cols = m + T;
    if (m <= WAVE) {
        if (cols <= 32) return launch_reg<1, 1, 32>(count, n, st, Q);
        if (cols <= 48) return launch_reg<2, 1, 24>(count, n, st, Q);
        if (cols <= 64) return launch_reg<2, 1, 32>(count, n, st, Q);
        return launch_reg<2, 1, 40>(count, n, st, Q);  // m + T <= 80
    }
    if (cols <= 96) return launch_reg<4, 2, 24>(count, n, st, Q);
    if (cols <= 128) return launch_reg<8, 2, 16>(count, n, st, Q);
    return launch_reg<8, 2, 24>(count, n, st, Q);  // m + T <= 144
}

// LDS bytes of the layout launch_f64_reg picks.
size_t f64_reg_lds(int m, int n, int T) {
    const int cols = m + T;
    if (m <= WAVE) return reg_lds_bytes(m, n, T, cols <= 32 ? 32 : cols <= 48 ? 48 : cols <= 64 ? 64 : 80, WAVE);
    return reg_lds_bytes(m, n, T, cols <= 96 ? 96 : cols <= 128 ? 128 : 192, 2 * WAVE);
}

// ---- persistent blocked LU on the matrix cores (FP64 default) -----------------------------
// One launch solves every stencil: LU_BLOCK-thread workgroups stay resident and take work
// units (one large stencil, or one stencil per wave or wave pair for small systems; largest
// systems first) from a device counter. A stencil's [M | R] lives in registers as 16x16 tiles
// in the accumulator layout of v_mfma_f64_16x16x4f64: register v of lane l of a tile holds
// row l%16, column 4v + l/16. Wave w of a W-wave team owns the column tiles w, w + W, ...
// (NCT of them); RT row tiles cover the m rows.
// Partial-pivot LU in panels of 8 columns (two K-steps of the instruction):
//   1. the owner wave of the panel's column tile copies the panel to LDS and factors it
//      row-per-lane: per column a pivot search over the unpivoted rows (the largest |a|,
//      ties to the smaller row, as in the unblocked kernels), multipliers, and the update of
//      the panel's remaining columns; the row order is kept in perm/pos instead of moving rows;
//   2. it stores the multipliers as the B operand L[r][q] (zero for rows pivoted earlier; a
//      row that became the pivot of column q keeps its multipliers of columns < q);
//   3. every wave forms the panel's eight U rows on its own columns (the pivot rows' values
//      with the unit-lower 8x8 block applied), zero on finished columns, as the A operand;
//   4. every wave updates its tiles with two v_mfma_f64_16x16x4f64 per (row tile, live
//      column tile): a[r][c] -= sum_q L[r][q] U[q][c], q = 0..3 then 4..7.
// The instruction computes D[i][j] = C[i][j] + sum_k A[i][k] B[k][j] with A[i][k] from lane
// i + 16k, B[k][j] from lane j + 16k and D[i][j] in register i/4 of lane j + 16(i%4), and
// accumulates the four products in order, each fused and rounded on its own (a probe of all
// 256 outputs on random operands equals four chained fma). Here i is the column and j the
// row: A[i][k] = -U[k][c], B[k][j] = L[r][k]. So every element sees the operations of the
// unblocked register kernel in the same order and the weights are bit-identical to it. The
// pivot rows' own updates in step 4 reproduce the U rows of step 3.
// After the last panel each lane stores its rows' U part and eliminated right-hand sides to
// LDS by pivot order, and back substitution runs as in the register kernel.
typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int LU_BLOCK = 256;   // threads per workgroup (4 waves)
constexpr int LU_WAVES = LU_BLOCK / WAVE;
constexpr int LU_NLAYOUT = 7;   // layouts 5 and 6 (128 and 192 doubles of tiles per lane) only in the wide kernel

struct LuLayout {
    int W, RT, NCT;
};

// W waves per stencil, RT row tiles, NCT column tiles per wave; a workgroup runs
// LU_WAVES / W stencils of one size at a time.
__host__ __device__ constexpr LuLayout lu_layout(int id) {
    return id == 0 ? LuLayout{1, 2, 2}
         : id == 1 ? LuLayout{2, 3, 2}
         : id == 2 ? LuLayout{2, 4, 2}
         : id == 3 ? LuLayout{4, 6, 2}
         : id == 4 ? LuLayout{4, 7, 2}   // m <= 112: the degree-2 systems of 101-node stencils
         : id == 5 ? LuLayout{4, 8, 2}
                   : LuLayout{4, 8, 3};
}

// Smallest layout holding m rows and m + T columns; -1 when none does.
__host__ __device__ inline int lu_pick(int m, int T) {
    for (int id = 0; id < LU_NLAYOUT; id++) {
        const LuLayout L = lu_layout(id);
        if (m <= 16 * L.RT && m + T <= 16 * L.W * L.NCT) return id;
    }
    return -1;
}

// Right-hand sides per wave in one back-substitution pass: LU_QC, or LU_QA when the
// right-hand-side buffer shares R0 with the packed U (see lu_rhs_in_r0).
constexpr int LU_QC = 4;
__host__ __device__ constexpr int lu_qa(int W) { return W == 1 ? 4 : W == 2 ? 2 : 1; }

// The right-hand-side buffer BS lives inside R0 (after the staged coordinates) when one
// back-substitution pass of LU_QA right-hand sides per wave covers all T: the waves then read
// their right-hand sides into registers before the packed U is written over them.
__host__ __device__ inline bool lu_rhs_in_r0(int W, int T) { return T <= W * lu_qa(W); }

// LDS of one team, in doubles: [R0][BS: m T, unless inside R0][dg: m][amax: 8][umax: 8], then
// [flag: 4 int] and 16-bit row numbers [perm: m][pos: 16 RT]. R0 holds the staged coordinates
// (5n) and the assembled right-hand sides during assembly, the panel, two multiplier buffers
// (each [8][16 RT + 4]) and the U-row buffer ([8][16 W NCT + 4]) during the factorisation, and
// the strict upper triangle by columns (m(m-1)/2) for back substitution. The panel buffers are
// column-major with a padded stride, so that the lanes of one column, and the four columns one
// wave reads at once, fall on different banks. With BS inside R0 and 16-bit row tables the
// largest degree-2 system (m = 111, 4 right-hand sides) takes 50.4 KB: three workgroups fit in
// the 160 KB of a CU (a workgroup's LDS is allocated in 2 KB steps).
__host__ __device__ inline size_t lu_r0(int W, int RT, int NCT, int m, int n, int T) {
    const size_t a = (size_t)5 * n + (lu_rhs_in_r0(W, T) ? (size_t)m * T : 0),
                 b = (size_t)24 * (16 * RT + 4) + (size_t)8 * (16 * W * NCT + 4), c = (size_t)m * (m - 1) / 2;
    const size_t ab = a > b ? a : b;
    return ab > c ? ab : c;
}
__host__ __device__ inline size_t lu_team_bytes(int W, int RT, int NCT, int m, int n, int T) {
    const size_t d = lu_r0(W, RT, NCT, m, n, T) + (lu_rhs_in_r0(W, T) ? 0 : (size_t)m * T) + m + 16;
    const size_t i = 4 * sizeof(int) + ((size_t)m + 16 * RT) * sizeof(short);
    return (d * sizeof(double) + i + 15) / 16 * 16;
}

// The value of x, hidden from the optimiser: lane-dependent indices derived from it are
// recomputed where they are used instead of being hoisted out of the persistent and panel
// loops, where dozens of them would stay live in VGPRs next to the matrix tiles.
__device__ __forceinline__ int opaque_v(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

template <int W>
__device__ __forceinline__ void team_sync() {
    if constexpr (W == 1)
        wave_lds_sync();
    else
        __syncthreads();
}

// Not inlined: each layout then gets its own register allocation (inlined into one kernel
// the layouts spill hundreds of VGPRs).
// Solves the stencils order[start .. start + count) (count <= number of teams), each of n
// nodes and m = n + np unknowns, one per team. One-wave teams without a stencil return at
// once; larger teams without one repeat the first team's solve and write nothing, so that
// every team passes the same workgroup barriers.
// OCC is the waves per SIMD of the calling kernel: the instances of the two kernels are kept
// apart so that each is compiled to its kernel's register budget.
// Arguments of a device function arrive in VGPRs, and values loaded through P are per lane to
// the compiler: the sizes, counts and flags are made wave-uniform (readfirstlane) on entry, so
// that the branches and loop bounds on them are scalar branches instead of exec-masked
// regions around every tile operation.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double uni(double x) {
    const unsigned long long u = __double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u), hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int W, int RT, int NCT, int OCC>
__device__ __attribute__((noinline)) void lu_team(const WParams& P, int start, int count, int n, int m, unsigned char* smem) {
    constexpr int MP = 16 * RT, NP = 16 * W * NCT, RR = (MP + WAVE - 1) / WAVE;
    constexpr int PS = MP + 4, US = NP + 4;  // strides of the panel / multiplier and U-row buffers
    constexpr int KB = 8;                    // panel width
    start = uni(start);
    count = uni(count);
    n = uni(n);
    m = uni(m);
    const int wv = uni((int)threadIdx.x / WAVE);
    int lane = opaque_v(threadIdx.x & (WAVE - 1));
    const int team = wv / W, wt = wv - team * W;
    if (W == 1 && team >= count) return;
    bool writer = team < count;
    const int s = uni(P.order[start + (writer ? team : 0)]);
    const int T = uni(P.T);
    const bool diag = uni(P.diag != nullptr ? 1 : 0) != 0;
    const int sph_basis = uni(P.sph);
    int rl = lane & 15, cg = lane >> 4;
    Sys S;
    S.P = &P;
    S.s = s;
    S.off = uni(P.st_off[s]);
    S.n = n;
    S.m = m;
    S.T = T;
    S.ld = 0;
    S.eps = uni(P.eps[s]);
    S.e2 = S.eps * S.eps;
    int sph = 0;
    for (int t = 0; t < T; t++) sph |= (P.types[t] == RBFX_LAMBDA || P.types[t] == RBFX_THETA);
    S.sph = uni(sph) != 0;
    S.XS = S.BS = nullptr;
    S.perm = nullptr;
    // team LDS
    unsigned char* base = smem + (size_t)team * lu_team_bytes(W, RT, NCT, m, n, T);
    const bool rhs_in_r0 = lu_rhs_in_r0(W, T);
    double* R0 = (double*)base;
    double* BS = rhs_in_r0 ? R0 + 5 * n : R0 + lu_r0(W, RT, NCT, m, n, T);
    double* dg = R0 + lu_r0(W, RT, NCT, m, n, T) + (rhs_in_r0 ? 0 : (size_t)m * T);
    double* amaxw = dg + m;
    double* umaxw = amaxw + 8;
    int* flag = (int*)(umaxw + 8);
    short* perm = (short*)(flag + 4);
    short* pos = perm + m;
    double* Pbuf = R0;                  // [KB][PS]
    double* Lbuf = R0 + KB * PS;        // [2][KB][PS]
    double* U12 = R0 + 3 * KB * PS;     // [KB][US]
    double* Upk = R0;                   // strict upper triangle of U by columns
    // ---- assembly: coordinates, right-hand sides, tiles ----
    S.X = R0;
    S.SPH = S.X + 3 * n;
    int tt = wt * WAVE + lane;
    for (int j = tt; j < n; j += WAVE * W) {
        const double* q = P.xyz + 3 * (size_t)P.st_idx[S.off + j];
        S.X[3 * j] = q[0];
        S.X[3 * j + 1] = q[1];
        S.X[3 * j + 2] = q[2];
        if (S.sph) {
            double lam, th;
            g_sph(q[0], q[1], q[2], lam, th);
            S.SPH[2 * j] = lam;
            S.SPH[2 * j + 1] = th;
        }
    }
    if (tt == 0) flag[0] = 0;
    for (int r = tt; r < MP; r += WAVE * W) pos[r] = (r < m) ? m : -1;  // -1: padding row, never a pivot
    team_sync<W>();
    for (int q = tt; q < m * T; q += WAVE * W) BS[q] = rhs_entry(S, q / T, q - (q / T) * T);
    team_sync<W>();
    d4v acc[RT][NCT];
    double amax = 0.0;
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
        const int r = 16 * rt + rl;
#pragma unroll
        for (int sl = 0; sl < NCT; sl++) {
            const int c0 = 16 * (wt + W * sl) + cg;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int c = c0 + 4 * v;
                double val = 0.0;
                if (r < m) {
                    if (c < m) {
                        val = m_entry(r, c, n, S.e2, S.X, sph_basis);
                        amax = fmax(amax, fabs(val));
                    } else if (c < m + T) {
                        val = BS[r * T + (c - m)];
                    }
                }
                acc[rt][sl][v] = val;
                __builtin_amdgcn_sched_barrier(0);  // one entry at a time: short live ranges
            }
        }
    }
    amax = wave_max(amax);
    if (lane == 0) amaxw[wt] = amax;
    team_sync<W>();  // the coordinates are read: R0 now holds the panel buffers
    // ---- blocked elimination, panels of KB = 8 columns (two K-steps of the instruction) ----
    const int RTL = (m + 15) >> 4;  // row tiles holding rows < m (the others are padding)
    for (int kp = 0; kp < m; kp += KB) {
        lane = opaque_v(lane);
        rl = lane & 15;
        cg = lane >> 4;
        const int npn = min(KB, m - kp);
        const int ctp = kp >> 4, owner = ctp % W, slot = ctp / W, gp = (kp & 15) >> 2;
        double* Lb = Lbuf + ((kp / KB) & 1) * (KB * PS);
        if (wt == owner) {
            // panel column kp + j sits in register gp + j/4 of lane group j%4
#pragma unroll
            for (int sl = 0; sl < NCT; sl++)
                if (sl == slot) {
#pragma unroll
                    for (int v = 0; v < 4; v += 2)
                        if (v == gp) {
#pragma unroll
                            for (int rt = 0; rt < RT; rt++) {
                                Pbuf[cg * PS + 16 * rt + rl] = acc[rt][sl][v];
                                Pbuf[(4 + cg) * PS + 16 * rt + rl] = acc[rt][sl][v + 1];
                            }
                        }
                }
            wave_lds_sync();
            double a[RR][KB];
            bool live[RR], was[RR];
            int newq[RR];
#pragma unroll
            for (int h = 0; h < RR; h++) {
                const int r = lane + WAVE * h;
                if (r < MP) {
#pragma unroll
                    for (int j = 0; j < KB; j++) a[h][j] = Pbuf[j * PS + r];
                    live[h] = pos[r] == m;
                } else {
#pragma unroll
                    for (int j = 0; j < KB; j++) a[h][j] = 0.0;
                    live[h] = false;
                }
                was[h] = live[h];
                newq[h] = KB;
            }
            bool bad = false;
#pragma unroll
            for (int q = 0; q < KB; q++) {
                if (q >= npn) break;  // uniform
                unsigned long long key = 0;
#pragma unroll
                for (int h = 0; h < RR; h++)
                    if (live[h]) {
                        const unsigned long long kh = pivot_key(a[h][q], lane + WAVE * h);
                        key = kh > key ? kh : key;
                    }
                key = wave_umax_uniform(key);
                const int p = 255 - (int)(key & 0xffull);
                const int pl = p & (WAVE - 1), ph = p / WAVE;
                const double piv = readlane_f64((RR == 1 || ph == 0) ? a[0][q] : a[RR - 1][q], pl);
                bad |= (key == 0) | (piv == 0.0);
#pragma unroll
                for (int h = 0; h < RR; h++)
                    if (lane + WAVE * h == p) {
                        live[h] = false;
                        newq[h] = q;
                        pos[p] = kp + q;
                    }
                if (lane == 0) {
                    perm[kp + q] = p;
                    dg[kp + q] = piv;
                }
                const double inv = 1.0 / piv;
                // the panel's later columns; in a last, partial panel the columns past m are
                // right-hand sides, which the matrix-core update below eliminates
                double pr[KB];
#pragma unroll
                for (int j = q + 1; j < KB; j++)
                    pr[j] = (j < npn) ? readlane_f64((RR == 1 || ph == 0) ? a[0][j] : a[RR - 1][j], pl) : 0.0;
#pragma unroll
                for (int h = 0; h < RR; h++)
                    if (live[h]) {
                        const double l = a[h][q] * inv;
                        a[h][q] = l;
#pragma unroll
                        for (int j = q + 1; j < KB; j++)
                            if (j < npn) a[h][j] = a[h][j] - l * pr[j];
                    }
            }
#pragma unroll
            for (int h = 0; h < RR; h++) {
                const int r = lane + WAVE * h;
                if (r < MP) {
#pragma unroll
                    for (int j = 0; j < KB; j++) {
                        Pbuf[j * PS + r] = a[h][j];
                        Lb[j * PS + r] = (was[h] && j < npn && j < newq[h]) ? a[h][j] : 0.0;
                    }
                }
            }
            if (lane == 0 && bad) flag[0] = 1;
            wave_lds_sync();
#pragma unroll
            for (int sl = 0; sl < NCT; sl++)
                if (sl == slot) {
#pragma unroll
                    for (int v = 0; v < 4; v += 2)
                        if (v == gp) {
#pragma unroll
                            for (int rt = 0; rt < RT; rt++) {
                                acc[rt][sl][v] = Pbuf[cg * PS + 16 * rt + rl];
                                acc[rt][sl][v + 1] = Pbuf[(4 + cg) * PS + 16 * rt + rl];
                            }
                        }
                }
        }
        team_sync<W>();  // multipliers, pivots and pos of this panel are visible
        const int kdone = kp + npn;  // columns < kdone are finished
        // 1. the pivot rows' values on this wave's live tiles: every lane checks whether its
        //    row of each row tile became the pivot of one of the panel's columns (pos), so the
        //    rows are found without a branch per pivot
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
            if (rt >= RTL) break;  // padding rows only (uniform)
            const int qq = (int)pos[16 * rt + rl] - kp;
            if (qq >= 0 && qq < npn) {
#pragma unroll
                for (int sl = 0; sl < NCT; sl++) {
                    const int ct = wt + W * sl;
                    if (16 * ct + 16 <= kdone || 16 * ct >= m + T) continue;
#pragma unroll
                    for (int v = 0; v < 4; v++) U12[qq * US + 16 * ct + 4 * v + cg] = acc[rt][sl][v];
                }
            }
        }
        wave_lds_sync();
        // 2. U rows: the unit-lower KB x KB block applied; lanes 16 sl .. 16 sl + 15 take the
        //    columns of tile slot sl
        if (lane < 16 * NCT) {
            const int ct = wt + W * (lane >> 4);
            const int c = 16 * ct + (lane & 15);
            if (16 * ct < m + T) {
                double u[KB];
                int pq[KB];
#pragma unroll
                for (int q = 0; q < KB; q++) {
                    u[q] = (q < npn) ? U12[q * US + c] : 0.0;
                    pq[q] = (q < npn) ? (int)perm[kp + q] : 0;
                }
#pragma unroll
                for (int q = 1; q < KB; q++) pq[q] = uni(pq[q]);  // one wait for all the pivot rows
#pragma unroll
                for (int q = 1; q < KB; q++)
                    if (q < npn) {
#pragma unroll
                        for (int jj = 0; jj < q; jj++) u[q] = u[q] - Lb[jj * PS + pq[q]] * u[jj];
                    }
                const bool fin = c < kdone;
#pragma unroll
                for (int q = 0; q < KB; q++) U12[q * US + c] = (fin || q >= npn) ? 0.0 : u[q];
            }
        }
        wave_lds_sync();
        // 3. trailing update, one instruction per K-step, row tile and live column tile
#pragma unroll
        for (int sl = 0; sl < NCT; sl++) {
            const int ct = wt + W * sl;
            if (16 * ct + 16 <= kdone || 16 * ct >= m + T) continue;  // finished or empty tile (uniform)
#pragma unroll
            for (int ks = 0; ks < KB / 4; ks++) {
                if (4 * ks >= npn) break;  // the U rows of this K-step are all zero (uniform)
                const double aop = -U12[(4 * ks + cg) * US + 16 * ct + rl];
#pragma unroll
                for (int rt = 0; rt < RT; rt++) {
                    if (rt >= RTL) break;  // padding rows only (uniform)
                    const double bop = Lb[(4 * ks + cg) * PS + 16 * rt + rl];
                    acc[rt][sl] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, bop, acc[rt][sl], 0, 0, 0);
                }
            }
        }
    }
    team_sync<W>();  // every tile is final; the panel buffers are free
    if (flag[0] != 0) {  // an exact zero pivot: NaN weights, the stencil flagged singular
        if (wt == 0 && writer) {
            for (int q = lane; q < n * T; q += WAVE) {
                const int j = q / T, t = q - j * T;
                P.w[(size_t)t * P.nnz + S.off + j] = NAN;
            }
            if (lane == 0) {
                P.fail_flags[s] = 1;
                if (diag) {
                    P.diag[4 * (size_t)s] = NAN;
                    P.diag[4 * (size_t)s + 1] = 0;
                    P.diag[4 * (size_t)s + 2] = 0;
                    P.diag[4 * (size_t)s + 3] = RBFX_PATH_SINGULAR;
                }
            }
        }
        if (W == 1 || LU_WAVES == W) return;  // the only team of its barriers
        writer = false;  // other teams share the workgroup barriers below: run on, write nothing
    }
    // U part (columns right of each row's pivot step) and eliminated right-hand sides by pivot
    // order. With BS inside R0 the right-hand sides go first: every wave reads its own into
    // registers before the packed U is written over them.
    lane = opaque_v(lane);
    rl = lane & 15;
    cg = lane >> 4;
    auto store_rows = [&](bool u_part, bool rhs_part) -> double {
        double um = 0.0;
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
            const int r = 16 * rt + rl;
            if (r >= m) continue;
            const int i = pos[r];
#pragma unroll
            for (int sl = 0; sl < NCT; sl++) {
                const int c0 = 16 * (wt + W * sl) + cg;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int c = c0 + 4 * v;
                    const double val = acc[rt][sl][v];
                    if (c < m) {
                        if (u_part) {
                            if (c >= i) um = fmax(um, fabs(val));
                            if (c > i) Upk[(size_t)c * (c - 1) / 2 + i] = val;
                        }
                    } else if (c < m + T) {
                        if (rhs_part) BS[i * T + (c - m)] = val;
                    }
                }
            }
        }
        return um;
    };
    // Back substitution of the right-hand sides t = t0 + wt + W q, q < Q: b[q][h] holds
    // unknown lane + 64 h (pivot order); then the weights out.
    auto load_rhs = [&](auto qtag, auto& b, int t0) {
        constexpr int Q = decltype(qtag)::value;
#pragma unroll
        for (int h = 0; h < RR; h++) {
            const int i = lane + WAVE * h;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int t = t0 + wt + W * q;
                b[q][h] = (t < T && i < m) ? BS[i * T + t] : 0.0;
            }
        }
    };
    auto back_sub = [&](auto qtag, auto& b, int t0) {
        constexpr int Q = decltype(qtag)::value;
        // column k of U (rows < k) and the pivot u_kk are loaded one step ahead
        double ucn[RR], ukn = dg[m - 1];
#pragma unroll
        for (int h = 0; h < RR; h++)
            ucn[h] = (lane + WAVE * h < m - 1) ? Upk[(size_t)(m - 1) * (m - 2) / 2 + lane + WAVE * h] : 0.0;
        for (int k = m - 1; k >= 0; k--) {
            double ucol[RR];
#pragma unroll
            for (int h = 0; h < RR; h++) ucol[h] = ucn[h];
            const double ukk = ukn;
            if (k > 0) {
                const double* uc = Upk + (size_t)(k - 1) * (k - 2) / 2;
#pragma unroll
                for (int h = 0; h < RR; h++) ucn[h] = (lane + WAVE * h < k - 1) ? uc[lane + WAVE * h] : 0.0;
                ukn = dg[k - 1];
            }
#pragma unroll
            for (int q = 0; q < Q; q++) {
                if (t0 + wt + W * q >= T) continue;
                const double bk = readlane_f64((RR == 1 || k < WAVE) ? b[q][0] : b[q][RR - 1], k & (WAVE - 1));
                const double x = bk / ukk;
#pragma unroll
                for (int h = 0; h < RR; h++) {
                    const int i = lane + WAVE * h;
                    if (i < k)
                        b[q][h] -= ucol[h] * x;
                    else if (i == k)
                        b[q][h] = x;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int t = t0 + wt + W * q;
            if (t >= T) continue;
            const double scale = (P.types[t] == RBFX_HV) ? P.hv_scale : 1.0;
#pragma unroll
            for (int h = 0; h < RR; h++) {
                const int i = lane + WAVE * h;
                if (i < n && writer) P.w[(size_t)t * P.nnz + S.off + i] = b[q][h] * scale;
            }
        }
    };
    // x by unknown into BS, for the diagnostics
    auto keep_x = [&](auto qtag, auto& b, int t0) {
        constexpr int Q = decltype(qtag)::value;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int t = t0 + wt + W * q;
            if (t >= T) continue;
#pragma unroll
            for (int h = 0; h < RR; h++) {
                const int i = lane + WAVE * h;
                if (i < m) BS[i * T + t] = b[q][h];
            }
        }
    };
    lane = opaque_v(lane);
    tt = wt * WAVE + lane;
    if (rhs_in_r0) {
        using QA = std::integral_constant<int, lu_qa(W)>;
        store_rows(false, true);
        team_sync<W>();
        double b[QA::value][RR];
        load_rhs(QA{}, b, 0);
        team_sync<W>();  // every wave holds its right-hand sides: BS may be overwritten
        const double um = wave_max(store_rows(true, false));
        if (lane == 0) umaxw[wt] = um;
        team_sync<W>();
        back_sub(QA{}, b, 0);
        if (diag) {
            team_sync<W>();  // every wave has read the packed U under BS
            keep_x(QA{}, b, 0);
        }
    } else {
        using QC = std::integral_constant<int, LU_QC>;
        const double um = wave_max(store_rows(true, true));
        if (lane == 0) umaxw[wt] = um;
        team_sync<W>();
        for (int t0 = 0; t0 < T; t0 += W * LU_QC) {
            double b[LU_QC][RR];
            load_rhs(QC{}, b, t0);
            back_sub(QC{}, b, t0);
            // each wave overwrites only the columns of BS it has read itself
            if (diag) keep_x(QC{}, b, t0);
        }
    }
    if (tt == 0 && writer) P.fail_flags[s] = 0;
    if (!diag) return;
    // diagnostics: x by unknown is in BS; the coordinates are restaged over the packed U
    team_sync<W>();  // back substitution has read the packed U
    for (int j = tt; j < n; j += WAVE * W) {
        const double* q = P.xyz + 3 * (size_t)P.st_idx[S.off + j];
        S.X[3 * j] = q[0];
        S.X[3 * j + 1] = q[1];
        S.X[3 * j + 2] = q[2];
        if (S.sph) {
            double lam, th;
            g_sph(q[0], q[1], q[2], lam, th);
            S.SPH[2 * j] = lam;
            S.SPH[2 * j + 1] = th;
        }
    }
    team_sync<W>();
    if (wt != 0 || !writer) return;
    double am = 0.0, umx = 0.0;
    for (int w = 0; w < W; w++) {
        am = fmax(am, amaxw[w]);
        umx = fmax(umx, umaxw[w]);
    }
    const double be = residual<double, MAX_T, OCC>(S, (double*)nullptr, BS);
    if (lane == 0) {
        P.diag[4 * (size_t)s] = be;
        P.diag[4 * (size_t)s + 1] = umx / am;
        P.diag[4 * (size_t)s + 2] = 0;
        P.diag[4 * (size_t)s + 3] = RBFX_PATH_F64;
    }
}

// The narrow kernel (layouts 0-4, tiles of at most 56 doubles per lane) is compiled for three
// waves per SIMD (168 VGPRs); the wide one (layouts 5 and 6: 64 and 96 doubles of tiles per
// lane) for two. With 64 doubles of tiles in 168 VGPRs the (4, 8, 2) layout spilled them; in the
// wide kernel the degree-3 set (m up to 121) takes 0.200 s instead of 0.209 s.
// Work units are numbered class by class (largest systems first): class c holds the units
// ustart[c] .. ustart[c + 1] - 1, and its record classes[c] = {first position in P.order,
// stencils, layout | n << 8, m}; unit k of the class solves the class's stencils k * teams ..
// k * teams + teams - 1 (teams = workgroup waves / layout waves).
template <bool WIDE>
__global__ void __launch_bounds__(LU_BLOCK) __attribute__((amdgpu_waves_per_eu(WIDE ? 2 : 3))) weights_lu_kernel(WParams P, const int4* __restrict__ classes, const int* __restrict__ ustart,
                                                             int n_cls, int n_units, int* __restrict__ counter) {
    constexpr int OCC = WIDE ? 2 : 3;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_unit;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(counter, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= n_units) return;  // the same value in every thread
        int c = 0;
        while (c + 1 < n_cls && u >= ustart[c + 1]) c++;
        const int4 C = classes[c];
        const int lay = C.z & 0xff, n = C.z >> 8, m = C.w;
        const int teams = LU_WAVES / lu_layout(lay).W;
        const int k = (u - ustart[c]) * teams;
        const int first = C.x + k, count = min(teams, C.y - k);
        switch (lay) {
            case 0: lu_team<1, 2, 2, OCC>(P, first, count, n, m, smem); break;
            case 1: lu_team<2, 3, 2, OCC>(P, first, count, n, m, smem); break;
            case 2: lu_team<2, 4, 2, OCC>(P, first, count, n, m, smem); break;
            case 3: lu_team<4, 6, 2, OCC>(P, first, count, n, m, smem); break;
            case 4: lu_team<4, 7, 2, OCC>(P, first, count, n, m, smem); break;
            case 5:
                if constexpr (WIDE) lu_team<4, 8, 2, OCC>(P, first, count, n, m, smem);
                break;
            default:
                if constexpr (WIDE) lu_team<4, 8, 3, OCC>(P, first, count, n, m, smem);
                break;
        }
        __syncthreads();  // the team LDS and s_unit are free for the next unit
    }
}

__global__ void zero_int_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// bad[0] = 1 and bad[1] = the index when some stencil index is outside [0, n_nodes).
__global__ void check_index_kernel(const int32_t* __restrict__ idx, int nnz, int n_nodes, int* bad) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nnz; q += (long long)gridDim.x * blockDim.x) {
        const int v = idx[q];
        if (v < 0 || v >= n_nodes) {
            atomicOr(bad, 1);
            bad[1] = v;
        }
    }
}

__global__ void mark_fallback_kernel(const int32_t* order, int count, double* diag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count && diag) diag[4 * (size_t)order[i] + 3] = RBFX_PATH_F32_FALLBACK_F64;
}


// Which FP64 solver runs: the persistent matrix-core LU (default), or, with RBFX_WEIGHTS_F64
// set to "reg" or "lds" (read at each call), the per-class register-resident or one-wave LDS
// kernels. All three give the same weights bit for bit.
enum class F64Solver { LU, REG, LDS };
F64Solver f64_solver() {
    const char* e = std::getenv("RBFX_WEIGHTS_F64");
    if (e && std::strcmp(e, "lds") == 0) return F64Solver::LDS;
    if (e && std::strcmp(e, "reg") == 0) return F64Solver::REG;
    return F64Solver::LU;
}

// Device scratch of one call: [order: n][flags: n + 2][classes: MAX_M + 1 int4][ustart:
// MAX_M + 2][counter]. 256-byte aligned pieces.
struct WScratch {
    int32_t* order;
    int* flags;
    int4* classes;
    int* ustart;
    int* counter;
};
constexpr size_t wscratch_al(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t wscratch_bytes(int n) {
    return wscratch_al(sizeof(int32_t) * (size_t)n) + wscratch_al(sizeof(int) * ((size_t)n + 2)) +
           wscratch_al(sizeof(int4) * (MAX_M + 1)) + wscratch_al(sizeof(int) * (MAX_M + 2)) + 256;
}
WScratch wscratch_carve(void* base, int n) {
    unsigned char* p = (unsigned char*)base;
    WScratch w;
    w.order = (int32_t*)p;
    p += wscratch_al(sizeof(int32_t) * (size_t)n);
    w.flags = (int*)p;
    p += wscratch_al(sizeof(int) * ((size_t)n + 2));
    w.classes = (int4*)p;
    p += wscratch_al(sizeof(int4) * (MAX_M + 1));
    w.ustart = (int*)p;
    p += wscratch_al(sizeof(int) * (MAX_M + 2));
    w.counter = (int*)p;
    return w;
}

// The whole weights computation on a caller-provided scratch of wscratch_bytes(n_stencils).
int weights_impl(rbfx_ctx* ctx, int32_t n_stencils, const int32_t* d_st_off, const int32_t* d_st_idx, int32_t n_nodes,
                 const double* d_xyz, const double* d_eps, uint32_t dertypes, int32_t dim, int32_t poly_deg, int32_t hv_k,
                 double hv_scale, int32_t mode, double* d_w, double* d_diag, int32_t* n_singular, void* d_scratch) {
    hipStream_t st = ctx->stream;
    WParams P;
    std::memset(&P, 0, sizeof(P));
    int T = 0;
    for (int b = 0; b < 16; b++)
        if (dertypes & (1u << b)) P.types[T++] = 1u << b;
    P.T = T;
    P.np = num_monomials(poly_deg);
    P.sph = (poly_deg >= 0 && (poly_deg & RBFX_POLY_SPHERE)) ? 1 : 0;
    P.dim = dim;
    P.hv_k = hv_k;
    P.hv_scale = hv_scale;
    P.st_off = d_st_off;
    P.st_idx = d_st_idx;
    P.xyz = d_xyz;
    P.eps = d_eps;
    P.w = d_w;
    P.diag = d_diag;
    P.max_iter = 8;

    // sizes -> classes (host side), validating the CSR offsets and stencil indices
    std::vector<int32_t> off(n_stencils + 1);
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(off.data(), d_st_off, sizeof(int32_t) * (n_stencils + 1), hipMemcpyDeviceToHost, st));
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(st));
    if (off[0] != 0) return rbfx_fail(ctx, RBFX_E_INVALID, "weights: st_off[0] = %d, expected 0", off[0]);
    int max_n = 0;
    for (int s = 0; s < n_stencils; s++) {
        int n = off[s + 1] - off[s];
        if (n < 1) return rbfx_fail(ctx, RBFX_E_INVALID, "weights: stencil %d has size %d", s, n);
        max_n = std::max(max_n, n);
    }
    if (max_n + num_monomials(effective_degree(poly_deg, max_n)) > MAX_M)
        return rbfx_fail(ctx, RBFX_E_TOO_LARGE, "weights: system size %d exceeds %d", max_n + num_monomials(effective_degree(poly_deg, max_n)), MAX_M);
    P.nnz = off[n_stencils];
    // counting sort of stencil ids by size, largest systems first
    std::vector<int> cnt(MAX_M + 1, 0);
    for (int s = 0; s < n_stencils; s++) cnt[off[s + 1] - off[s]]++;
    std::vector<int> start(MAX_M + 2, 0);
    for (int n = MAX_M; n >= 1; n--) start[n - 1] = start[n] + cnt[n];  // start[n-1]: first slot of size n
    std::vector<int32_t> order(n_stencils);
    {
        std::vector<int> fill(MAX_M + 1);
        for (int n = 1; n <= MAX_M; n++) fill[n] = start[n - 1] - cnt[n];
        for (int s = 0; s < n_stencils; s++) {
            int n = off[s + 1] - off[s];
            order[fill[n]++] = s;
        }
    }
    const WScratch W = wscratch_carve(d_scratch, n_stencils);
    int32_t* d_order = W.order;
    int* d_flags = W.flags;
    hipError_t e;
    {  // every stencil index must name a node: checked on the device before any solve reads it
        int* d_bad = d_flags + n_stencils;
        int bad[2] = {0, 0};
        const int blocks = (int)std::min<long long>(4096, ((long long)P.nnz + 255) / 256);
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, st, d_bad, 2);
        hipLaunchKernelGGL(check_index_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, st, d_st_idx, P.nnz, n_nodes, d_bad);
        if ((e = hipGetLastError()) == hipSuccess &&
            (e = hipMemcpyAsync(bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st)) == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e != hipSuccess) return rbfx_fail(ctx, RBFX_E_HIP, "weights: %s", hipGetErrorString(e));
        if (bad[0]) return rbfx_fail(ctx, RBFX_E_INVALID, "weights: stencil index %d out of range", bad[1]);
    }
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(d_order, order.data(), sizeof(int32_t) * n_stencils, hipMemcpyHostToDevice, st));
    P.fail_flags = d_flags;
    P.order = d_order;
    int lds_attr = 0;
    if (hipDeviceGetAttribute(&lds_attr, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess || lds_attr <= 0)
        lds_attr = 64 * 1024;
    const size_t lds_limit = (size_t)lds_attr;
    const F64Solver solver = f64_solver();

    // One persistent launch of the matrix-core LU over the classes sizes_count / sizes_start of
    // the order array (largest systems first). The class table (per class its stencils and
    // layout, and the number of the class's first work unit) goes to the scratch; the kernel
    // derives each unit from it.
    std::vector<int4> classes;
    std::vector<int> ustart;
    auto launch_persistent = [&](const std::vector<int>& sizes_count, const std::vector<int>& sizes_start) -> int {
        classes.clear();
        ustart.assign(1, 0);
        size_t lds = 0;
        bool wide = false;
        for (int n = MAX_M; n >= 1; n--) {
            const int c = sizes_count[n];
            if (!c) continue;
            const int m = n + num_monomials(effective_degree(poly_deg, n));
            const int id = lu_pick(m, T);
            if (id < 0) return rbfx_fail(ctx, RBFX_E_TOO_LARGE, "weights: system %d x %d with %d RHS has no layout", m, m, T);
            const LuLayout L = lu_layout(id);
            wide |= id >= 5;
            const int teams = LU_WAVES / L.W;
            lds = std::max(lds, (size_t)teams * lu_team_bytes(L.W, L.RT, L.NCT, m, n, T));
            classes.push_back(make_int4(sizes_start[n], c, id | (n << 8), m));
            ustart.push_back(ustart.back() + (c + teams - 1) / teams);
        }
        const int n_units = ustart.back();
        if (n_units == 0) return RBFX_OK;
        if (lds + STATIC_LDS_RESERVE > lds_limit)
            return rbfx_fail(ctx, RBFX_E_TOO_LARGE, "weights: matrix-core LU needs %zu B of LDS", lds);
        hipError_t le;
        if ((le = hipMemcpyAsync(W.classes, classes.data(), sizeof(int4) * classes.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
            (le = hipMemcpyAsync(W.ustart, ustart.data(), sizeof(int) * ustart.size(), hipMemcpyHostToDevice, st)) != hipSuccess)
            return rbfx_fail(ctx, RBFX_E_HIP, "weights: %s", hipGetErrorString(le));
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, st, W.counter, 1);
        int per_cu = 0;
        le = wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, weights_lu_kernel<true>, LU_BLOCK, lds)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, weights_lu_kernel<false>, LU_BLOCK, lds);
        if (le != hipSuccess || per_cu < 1)
            return rbfx_fail(ctx, RBFX_E_HIP, "weights: no occupancy for the matrix-core LU (%zu B LDS)", lds);
        const int grid = (int)std::min<long long>((long long)n_units, (long long)per_cu * std::max(ctx->n_cu, 1));
        WParams Q = P;
        Q.fp32 = 0;
        const int n_cls = (int)classes.size();
        if (wide)
            hipLaunchKernelGGL(weights_lu_kernel<true>, dim3(grid), dim3(LU_BLOCK), lds, st, Q, (const int4*)W.classes,
                               (const int*)W.ustart, n_cls, n_units, W.counter);
        else
            hipLaunchKernelGGL(weights_lu_kernel<false>, dim3(grid), dim3(LU_BLOCK), lds, st, Q, (const int4*)W.classes,
                               (const int*)W.ustart, n_cls, n_units, W.counter);
        le = hipGetLastError();
        if (le != hipSuccess) return rbfx_fail(ctx, RBFX_E_HIP, "weights: matrix-core LU launch failed: %s", hipGetErrorString(le));
        return RBFX_OK;
    };

    // One launch per size class (the FP32 refinement kernel, and the FP64 register / LDS kernels).
    auto launch_classes = [&](bool ir, const std::vector<int>& sizes_count, const std::vector<int>& sizes_start) -> int {
        for (int n = MAX_M; n >= 1; n--) {
            int c = sizes_count[n];
            if (!c) continue;
            WParams Q = P;
            Q.order = d_order + sizes_start[n];
            Q.count = c;
            Q.np = num_monomials(effective_degree(poly_deg, n));
            Q.m = n + Q.np;
            Q.fp32 = ir ? 1 : 0;
            const bool small_t = T <= 4;
            if (!ir && solver == F64Solver::REG) {
                const size_t rl = f64_reg_lds(Q.m, n, T);
                if (rl + STATIC_LDS_RESERVE > lds_limit)
                    return rbfx_fail(ctx, RBFX_E_TOO_LARGE, "weights: system %d x %d with %d RHS needs %zu B of LDS", Q.m, Q.m, T, rl);
                hipError_t le = launch_f64_reg(Q.m, T, c, n, st, Q);
                if (le != hipSuccess) return rbfx_fail(ctx, RBFX_E_HIP, "weights: launch (m=%d) failed: %s", Q.m, hipGetErrorString(le));
                continue;
            }
            size_t lds = ir ? lds_bytes<float>(Q.m, n, T, true) : lds_bytes<double>(Q.m, n, T, false);
            if (lds + STATIC_LDS_RESERVE > lds_limit) return rbfx_fail(ctx, RBFX_E_TOO_LARGE, "weights: system %d x %d with %d RHS needs %zu B of LDS", Q.m, Q.m, T, lds);
            if (ir && small_t)
                hipLaunchKernelGGL(weights_f32ir_kernel<4>, dim3(c), dim3(WAVE), lds, st, Q);
            else if (ir)
                hipLaunchKernelGGL(weights_f32ir_kernel<MAX_T>, dim3(c), dim3(WAVE), lds, st, Q);
            else if (small_t)
                hipLaunchKernelGGL(weights_f64_kernel<4>, dim3(c), dim3(WAVE), lds, st, Q);
            else
                hipLaunchKernelGGL(weights_f64_kernel<MAX_T>, dim3(c), dim3(WAVE), lds, st, Q);
            hipError_t le = hipGetLastError();
            if (le != hipSuccess) return rbfx_fail(ctx, RBFX_E_HIP, "weights: launch (m=%d) failed: %s", Q.m, hipGetErrorString(le));
        }
        return RBFX_OK;
    };
    auto launch_f64 = [&](const std::vector<int>& sizes_count, const std::vector<int>& sizes_start) -> int {
        return solver == F64Solver::LU ? launch_persistent(sizes_count, sizes_start)
                                       : launch_classes(false, sizes_count, sizes_start);
    };
    std::vector<int> cstart(MAX_M + 1, 0);
    for (int n = 1; n <= MAX_M; n++) cstart[n] = start[n - 1] - cnt[n];
    int rc = (mode == RBFX_SOLVE_F32_LU_F64_IR) ? launch_classes(true, cnt, cstart) : launch_f64(cnt, cstart);
    if (rc) {
        hipStreamSynchronize(st);
        return rc;
    }

    std::vector<int> flags(n_stencils);
    if ((e = hipMemcpyAsync(flags.data(), d_flags, sizeof(int) * n_stencils, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return rbfx_fail(ctx, RBFX_E_HIP, "weights: %s", hipGetErrorString(e));
    if (mode == RBFX_SOLVE_F32_LU_F64_IR) {
        // FP64 re-solve of the stencils whose refinement did not converge
        std::vector<int32_t> redo;
        std::vector<int> rcnt(MAX_M + 1, 0), rstart(MAX_M + 1, 0);
        for (int n = MAX_M; n >= 1; n--)
            for (int q = cstart[n]; q < cstart[n] + cnt[n]; q++)
                if (flags[order[q]] == 2) { redo.push_back(order[q]); rcnt[n]++; }
        if (!redo.empty()) {
            int acc = 0;
            for (int n = MAX_M; n >= 1; n--) { rstart[n] = acc; acc += rcnt[n]; }
            RBFX_HIP_CHECK(ctx, hipMemcpyAsync(d_order, redo.data(), sizeof(int32_t) * redo.size(), hipMemcpyHostToDevice, st));
            rc = launch_f64(rcnt, rstart);
            if (rc) {
                hipStreamSynchronize(st);
                return rc;
            }
            if (d_diag)
                hipLaunchKernelGGL(mark_fallback_kernel, dim3((unsigned)((redo.size() + 255) / 256)), dim3(256), 0, st, d_order,
                                   (int)redo.size(), d_diag);
            std::vector<int> f2(n_stencils);
            if ((e = hipMemcpyAsync(f2.data(), d_flags, sizeof(int) * n_stencils, hipMemcpyDeviceToHost, st)) != hipSuccess ||
                (e = hipStreamSynchronize(st)) != hipSuccess)
                return rbfx_fail(ctx, RBFX_E_HIP, "weights: %s", hipGetErrorString(e));
            for (int32_t s : redo) flags[s] = f2[s];
        }
    }
    int sing = 0;
    for (int s = 0; s < n_stencils; s++) sing += (flags[s] == 1);
    if (n_singular) *n_singular = sing;
    if (sing) return rbfx_fail(ctx, RBFX_E_SINGULAR, "weights: %d stencil system(s) singular", sing);
    return RBFX_OK;
}

// Argument checks shared by the entry points; RBFX_OK when the call may proceed.
int weights_check(rbfx_ctx* ctx, int32_t n_stencils, const int32_t* d_st_off, const int32_t* d_st_idx, int32_t n_nodes,
                  const double* d_xyz, const double* d_eps, uint32_t dertypes, int32_t dim, int32_t poly_deg, int32_t hv_k,
                  int32_t mode, double* d_w) {
    if (n_stencils <= 0 || !d_st_off || !d_st_idx || n_nodes <= 0 || !d_xyz || !d_eps || !d_w)
        return rbfx_fail(ctx, RBFX_E_INVALID, "weights: bad arguments");
    const uint32_t supported = RBFX_X | RBFX_Y | RBFX_Z | RBFX_LAPL | RBFX_R | RBFX_HV | RBFX_LAMBDA | RBFX_THETA |
                               RBFX_LSFC | RBFX_XSFC | RBFX_YSFC | RBFX_ZSFC | RBFX_INTERP;
    if (!dertypes || (dertypes & ~supported))
        return rbfx_fail(ctx, RBFX_E_INVALID, "weights: unsupported derivative mask 0x%x", dertypes);
    if (dim < 1 || dim > 3) return rbfx_fail(ctx, RBFX_E_INVALID, "weights: dim %d not in 1..3", dim);
    if (poly_deg < -1 || (poly_deg >= 0 && ((poly_deg & 15) > 3 || (poly_deg & ~(15 | RBFX_POLY_SPHERE)))))
        return rbfx_fail(ctx, RBFX_E_INVALID, "weights: poly_deg %d not in -1..3 (| RBFX_POLY_SPHERE)", poly_deg);
    if (hv_k < 0 || hv_k > 16) return rbfx_fail(ctx, RBFX_E_INVALID, "weights: hv_k %d", hv_k);
    if (mode != RBFX_SOLVE_F64_LU && mode != RBFX_SOLVE_F32_LU_F64_IR) return rbfx_fail(ctx, RBFX_E_INVALID, "weights: mode %d", mode);
    return RBFX_OK;
}

}  // namespace

// Scratch of rbfx_weights_noalloc: one copy per device, at a per-device address taken once with
// hipGetSymbolAddress (the symbol itself is a host-side stand-in, not a device address).
static __device__ __align__(256) unsigned char scratch[wscratch_bytes(RBFX_WEIGHTS_NOALLOC_MAX_STENCILS)];
namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of this device's scratch array (ctx's device current), found on the first
// call on the device.
int noalloc_scratch(rbfx_ctx* ctx, void** out) {
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
        return rbfx_fail(ctx, RBFX_E_INVALID, "weights_noalloc: device %d beyond %d", ctx->device, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return rbfx_fail(ctx, RBFX_E_HIP, "weights_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_scratch_ptr[ctx->device] = ptr;
    }
    *out = g_scratch_ptr[ctx->device];
    return RBFX_OK;
}
}  // namespace

extern "C" int rbfx_weights_noalloc(rbfx_ctx* ctx, int32_t n_stencils, const int32_t* d_st_off, const int32_t* d_st_idx,
                                    int32_t n_nodes, const double* d_xyz, const double* d_eps, uint32_t dertypes,
                                    int32_t dim, int32_t poly_deg, int32_t hv_k, double hv_scale, int32_t mode,
                                    double* d_w, double* d_diag, int32_t* n_singular) {
    if (!ctx) return RBFX_E_INVALID;
    if (n_singular) *n_singular = 0;
    int rc = weights_check(ctx, n_stencils, d_st_off, d_st_idx, n_nodes, d_xyz, d_eps, dertypes, dim, poly_deg, hv_k, mode, d_w);
    if (rc) return rc;
    if (n_stencils > RBFX_WEIGHTS_NOALLOC_MAX_STENCILS)
        return rbfx_fail(ctx, RBFX_E_TOO_LARGE, "weights_noalloc: %d stencils exceed the scratch maximum of %d", n_stencils,
                         RBFX_WEIGHTS_NOALLOC_MAX_STENCILS);
    RBFX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    void* d_scratch = nullptr;
    if ((rc = noalloc_scratch(ctx, &d_scratch))) return rc;
    rc = weights_impl(ctx, n_stencils, d_st_off, d_st_idx, n_nodes, d_xyz, d_eps, dertypes, dim, poly_deg, hv_k, hv_scale, mode,
                      d_w, d_diag, n_singular, d_scratch);
    hipStreamSynchronize(ctx->stream);
    return rc;
}

extern "C" int rbfx_weights(rbfx_ctx* ctx, int32_t n_stencils, const int32_t* d_st_off, const int32_t* d_st_idx,
                            int32_t n_nodes, const double* d_xyz, const double* d_eps, uint32_t dertypes, int32_t dim,
                            int32_t poly_deg, int32_t hv_k, double hv_scale, int32_t mode, double* d_w, double* d_diag,
                            int32_t* n_singular) {
    if (!ctx) return RBFX_E_INVALID;
    if (n_singular) *n_singular = 0;
    int rc = weights_check(ctx, n_stencils, d_st_off, d_st_idx, n_nodes, d_xyz, d_eps, dertypes, dim, poly_deg, hv_k, mode, d_w);
    if (rc) return rc;
    RBFX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    void* d_scratch = nullptr;
    hipError_t e = hipMalloc(&d_scratch, wscratch_bytes(n_stencils));
    if (e != hipSuccess) return rbfx_fail(ctx, RBFX_E_NOMEM, "weights: %s", hipGetErrorString(e));
    rc = weights_impl(ctx, n_stencils, d_st_off, d_st_idx, n_nodes, d_xyz, d_eps, dertypes, dim, poly_deg, hv_k, hv_scale, mode,
                      d_w, d_diag, n_singular, d_scratch);
    hipStreamSynchronize(ctx->stream);
    hipFree(d_scratch);
    return rc;
}
