// CPU restatement of the rbffd_gpu hot path. TEST INFRASTRUCTURE ONLY: tests/, bench.py's
// cpu_baseline leg and __graft_entry__.smoke() use it as the checker; the product never
// links, loads or calls it.
//
// Built with -ffp-contract=off so every sum and product rounds where the reference's code
// rounds (no fused multiply-adds), and with operations in the reference's order.
//
// Pinning (what checks this restatement independently of its own formulas):
//   Gaussian primitives (rbf_gaussian.h:40-249)     central finite differences of eval;
//        Laguerre closed form of p_k for dim=2 (scipy.special.eval_laguerre);
//        explicit k=1..3 polynomials from rbf_gaussian.h:204-228 (the #if 0 block)
//                                                     -> tests/test_oracle.py
//   LHS/RHS + solve (rbffd.cpp:129-191,225-352,356-537,888-969)
//        scipy.linalg.solve (LAPACK getrf/getrs, the routine arma::solve reaches) on the
//        same systems; polynomial reproduction; Laplace-Beltrami and surface gradient of
//        spherical harmonics vs closed forms; 5-point flat-limit Laplacian KAT
//        (tests/unified/compute_weights.cxx:313-315)   -> tests/test_oracle.py
//   apply / SpMM (rbffd.cpp:669-688)                  exact small-integer KAT, brute force
//   RK4 (time_dependent_pde.cpp:172-304)              amplification factor of du/dt = lambda u
//   cosine bell (cosine_bell.cpp:23-70, exact_advection.h:37-54, pde.cpp:463-533)
//        RHS of a smooth field vs the analytic -u.grad h; exact bell closed form at the
//        centre/edge; norms vs numpy
//   weights with poly_deg >= 1 are this build's addition (the reference only has the
//   constant monomial); they are pinned by polynomial reproduction, not by the reference.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---- Gaussian RBF (src/rbffd/rbfs/rbf_gaussian.h) --------------------------------------
struct Gauss {
    double eps, eps2;
    int dim;
    Gauss(double e, int d) : eps(e), eps2(e * e), dim(d) {}  // rbf.h:32-36
    static double square(double x, double y, double z) { return x * x + y * y + z * z; }  // Vec3d.h:104-106
    double eval(double x, double y, double z) const {  // rbf_gaussian.h:40-43
        double r2 = square(x, y, z);
        return std::exp(-(r2 * eps2));
    }
    double xderiv(double x, double y, double z) const {  // rbf_gaussian.h:61-65
        double xeps = x * eps2;
        return -2. * xeps * eval(x, y, z);
    }
    double yderiv(double x, double y, double z) const {  // rbf_gaussian.h:73-77
        double yeps = y * eps2;
        return -2. * yeps * eval(x, y, z);
    }
    double zderiv(double x, double y, double z) const {  // rbf_gaussian.h:85-89
        double zeps = z * eps2;
        return -2. * zeps * eval(x, y, z);
    }
    double rderiv(double x, double y, double z) const {  // rbf_gaussian.h:98-102
        double r2 = square(x, y, z);
        double reps = std::sqrt(r2) * eps2;
        return -2. * reps * eval(x, y, z);
    }
    double rderiv_over_r(double x, double y, double z) const {  // rbf_gaussian.h:112-115
        return -2. * eps2 * eval(x, y, z);
    }
    double rderiv2(double x, double y, double z) const {  // rbf_gaussian.h:124-127
        double r2 = square(x, y, z);
        return rderiv_over_r(x, y, z) + 4. * (eps2 * eps2) * r2 * std::exp(-(eps2 * r2));
    }
    double lapl(double x, double y, double z) const {  // rbf.h:304-318 dispatch
        switch (dim) {
            case 1: {  // rbf_gaussian.h:138-141
                double x2eps2 = x * x * eps2;
                return 2. * eps2 * (-1. + 2. * x2eps2) * eval(x, y, z);
            }
            case 2: {  // rbf_gaussian.h:148-152
                double x2eps2 = x * x * eps2;
                double y2eps2 = y * y * eps2;
                return 4. * eps2 * (-1. + x2eps2 + y2eps2) * eval(x, y, z);
            }
            case 3: {  // rbf_gaussian.h:160-164
                double r2 = square(x, y, z);
                double r2eps4 = r2 * (eps2 * eps2);
                return (-6. * eps2 + 4. * r2eps4) * eval(x, y, z);
            }
            default:
                return NAN;
        }
    }
    double hv_p_k(double r2, int k) const {  // rbf_gaussian.h:232-249 (recursive as written)
        double eps2r2 = eps2 * r2;
        switch (k) {
            case 0: return 1.;
            case 1: return 4. * (eps2r2)-2. * dim;
            default: break;
        }
        return 4. * (eps2r2 - 2. * (k - 1) - dim / 2.) * hv_p_k(r2, k - 1) -
               8. * (k - 1) * (2. * (k - 1) - 2. + dim) * hv_p_k(r2, k - 2);
    }
    double hyperviscosity(double x, double y, double z, int k) const {  // rbf_gaussian.h:200-229
        double r2 = square(x, y, z);
        return std::pow(eps, 2 * k) * hv_p_k(r2, k) * eval(x, y, z);
    }
};

// ---- spherical coordinates (src/utils/geom/cart2sph.cpp:3-11) --------------------------
struct Sph {
    double theta, phi, r;
};
inline Sph cart2sph(double x, double y, double z) {
    Sph s;
    s.theta = std::atan2(y, x);
    s.r = std::sqrt(x * x + y * y + z * z);
    s.phi = std::atan2(z, std::sqrt(x * x + y * y));
    return s;
}

// ---- polynomial basis (this build's addition for poly_deg >= 1) ------------------------
// Monomials of total degree 0..deg in centred coordinates d = x_j - x_centre, ordered by
// degree, then by descending power of x, then of y.
struct Mono {
    int a, b, c;
};
// poly_deg carries the degree in its low 4 bits (-1 = none) and, in bit 4 (value 16), the
// sphere-reduced basis: on the unit sphere |d|^2 + 2 x0.d = 0, so dz^2 is dropped at degree 2
// and dz^2, x dz^2, y dz^2, dz^3 at degree 3 (9 and 16 functions, the dimensions of the
// spherical harmonics of degree <= 2 and <= 3).
const int kPolySphere = 16;
inline int mono_list(int poly_deg, Mono* out) {
    if (poly_deg < 0) return 0;
    int deg = poly_deg & 15;
    bool sph = (poly_deg & kPolySphere) != 0;
    int cnt = 0;
    for (int d = 0; d <= deg; d++)
        for (int a = d; a >= 0; a--)
            for (int b = d - a; b >= 0; b--) {
                int c = d - a - b;
                if (sph && c >= 2) continue;  // (0,0,2) (1,0,2) (0,1,2) (0,0,3)
                out[cnt++] = {a, b, c};
            }
    return cnt;
}
inline double mono_eval(const Mono& m, double dx, double dy, double dz) {
    double v = 1.0;
    for (int i = 0; i < m.a; i++) v *= dx;
    for (int i = 0; i < m.b; i++) v *= dy;
    for (int i = 0; i < m.c; i++) v *= dz;
    return v;
}

// Value of operator `type` applied to monomial m at the stencil centre x0 (where d = 0).
// Only the constant, linear and quadratic terms have nonzero value, gradient or Hessian at d = 0.
double mono_rhs(uint32_t type, const Mono& m, const double* x0, int dim, int hv_k) {
    int deg = m.a + m.b + m.c;
    double p0 = (deg == 0) ? 1.0 : 0.0;
    double g[3] = {0, 0, 0};
    double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (deg == 1) {
        g[0] = m.a; g[1] = m.b; g[2] = m.c;
    } else if (deg == 2) {
        int e[3] = {m.a, m.b, m.c};
        for (int i = 0; i < 3; i++) {
            if (e[i] == 2) H[i][i] = 2.0;
            for (int j = i + 1; j < 3; j++)
                if (e[i] == 1 && e[j] == 1) H[i][j] = H[j][i] = 1.0;
        }
    }
    double tr = 0;
    for (int i = 0; i < dim && i < 3; i++) tr += H[i][i];
    switch (type) {
        case 0x1: return g[0];
        case 0x2: return g[1];
        case 0x4: return g[2];
        case 0x8: return tr;
        case 0x10: return 0.0;
        case 0x20: return (hv_k == 0) ? p0 : (hv_k == 1 ? tr : 0.0);
        case 0x40: {  // (1/cos theta) d/dlambda on the unit sphere at x0's angles
            Sph s = cart2sph(x0[0], x0[1], x0[2]);
            return -std::sin(s.theta) * g[0] + std::cos(s.theta) * g[1];
        }
        case 0x80: {  // d/dtheta (latitude)
            Sph s = cart2sph(x0[0], x0[1], x0[2]);
            return -std::sin(s.phi) * std::cos(s.theta) * g[0] - std::sin(s.phi) * std::sin(s.theta) * g[1] +
                   std::cos(s.phi) * g[2];
        }
        case 0x100: {  // Laplace-Beltrami: tr(H) - x0' H x0 - 2 x0.g
            double full = H[0][0] + H[1][1] + H[2][2];
            double xhx = 0, xg = 0;
            for (int i = 0; i < 3; i++) {
                xg += x0[i] * g[i];
                for (int j = 0; j < 3; j++) xhx += x0[i] * H[i][j] * x0[j];
            }
            return full - xhx - 2.0 * xg;
        }
        case 0x200: case 0x400: case 0x800: {  // surface gradient component (I - x0 x0') g
            int c = (type == 0x200) ? 0 : (type == 0x400 ? 1 : 2);
            double xg = x0[0] * g[0] + x0[1] * g[1] + x0[2] * g[2];
            return g[c] - x0[c] * xg;
        }
        case 0x8000: return p0;
        default: return 0.0;  // ALT_* (rbffd.cpp:490-494)
    }
}

// RHS entry of RBF j for operator `type` (rbffd.cpp:356-537).
double rbf_rhs(uint32_t type, const Gauss& rbf, const double* x0v, const double* xjv, int hv_k) {
    double dx = x0v[0] - xjv[0], dy = x0v[1] - xjv[1], dz = x0v[2] - xjv[2];  // diff = x0v - xjv (399)
    switch (type) {
        case 0x1: return rbf.xderiv(dx, dy, dz);
        case 0x2: return rbf.yderiv(dx, dy, dz);
        case 0x4: return rbf.zderiv(dx, dy, dz);
        case 0x8: return rbf.lapl(dx, dy, dz);
        case 0x8000: return rbf.eval(dx, dy, dz);
        case 0x10: return rbf.rderiv(dx, dy, dz);
        case 0x40: {  // rbffd.cpp:433-446
            Sph sj = cart2sph(xjv[0], xjv[1], xjv[2]);
            Sph si = cart2sph(x0v[0], x0v[1], x0v[2]);
            double dr_dlambda = std::cos(sj.phi) * std::sin(si.theta - sj.theta);
            return (std::fabs(dr_dlambda) > 0.) ? dr_dlambda * rbf.rderiv_over_r(dx, dy, dz) : 0.;
        }
        case 0x80: {  // rbffd.cpp:447-457
            Sph sj = cart2sph(xjv[0], xjv[1], xjv[2]);
            Sph si = cart2sph(x0v[0], x0v[1], x0v[2]);
            double dr_dtheta = std::cos(sj.phi) * std::sin(si.phi) * std::cos(si.theta - sj.theta) -
                               std::sin(sj.phi) * std::cos(si.phi);
            return (std::fabs(dr_dtheta) > 0.) ? dr_dtheta * rbf.rderiv_over_r(dx, dy, dz) : 0.;
        }
        case 0x100: {  // rbffd.cpp:458-463
            double r2 = Gauss::square(dx, dy, dz);
            return (1. / 4.) * ((4. - r2) * rbf.rderiv2(dx, dy, dz) + (4. - 3. * r2) * rbf.rderiv_over_r(dx, dy, dz));
        }
        case 0x200: case 0x400: case 0x800: {  // rbffd.cpp:465-489
            double xTx_k = x0v[0] * xjv[0] + x0v[1] * xjv[1] + x0v[2] * xjv[2];
            int c = (type == 0x200) ? 0 : (type == 0x400 ? 1 : 2);
            double v = x0v[c] * xTx_k - xjv[c];
            return v * rbf.rderiv_over_r(dx, dy, dz);
        }
        case 0x1000: case 0x2000: case 0x4000: return 0.;
        case 0x20: return rbf.hyperviscosity(dx, dy, dz, hv_k);  // rbffd.cpp:495-498
        default: return NAN;
    }
}

int num_types(uint32_t m) { return __builtin_popcount(m & 0xFFFFu); }

}  // namespace

extern "C" {

double orc_rbf_eval(double eps, double dx, double dy, double dz) { return Gauss(eps, 3).eval(dx, dy, dz); }
double orc_rbf_xderiv(double eps, double dx, double dy, double dz) { return Gauss(eps, 3).xderiv(dx, dy, dz); }
double orc_rbf_lapl(double eps, int dim, double dx, double dy, double dz) { return Gauss(eps, dim).lapl(dx, dy, dz); }
double orc_rbf_rderiv_over_r(double eps, double dx, double dy, double dz) {
    return Gauss(eps, 3).rderiv_over_r(dx, dy, dz);
}
double orc_rbf_rderiv2(double eps, double dx, double dy, double dz) { return Gauss(eps, 3).rderiv2(dx, dy, dz); }
double orc_rbf_hv(double eps, int dim, int k, double dx, double dy, double dz) {
    return Gauss(eps, dim).hyperviscosity(dx, dy, dz, k);
}
double orc_hv_pk(double eps, int dim, int k, double r2) { return Gauss(eps, dim).hv_p_k(r2, k); }

int orc_num_monomials(int poly_deg) {
    if (poly_deg < 0) return 0;
    if ((poly_deg & 15) > 3 || (poly_deg & ~31)) return -1;
    Mono tmp[20];
    return mono_list(poly_deg, tmp);
}

// Degree actually used for a stencil of n nodes: the largest d <= poly_deg whose monomial
// count does not exceed n (more monomials than nodes make [A P; P' 0] singular, e.g. degree 3
// needs 20 nodes, so an n = 17 stencil falls back to degree 2). This build's rule.
int orc_effective_degree(int poly_deg, int n) {
    if (poly_deg < 0) return poly_deg;
    int flag = poly_deg & kPolySphere, d = poly_deg & 15;
    while (d > 0 && orc_num_monomials(d | flag) > n) d--;
    return d | flag;
}

// getStencilLHS + distanceMatrix (rbffd.cpp:129-191, 888-969) and getStencilMultiRHS
// (rbffd.cpp:106-125). Row-major M[r*m + c], R[r*T + t].
int orc_stencil_system(const double* xyz, const int32_t* st, int32_t n, double eps, uint32_t dertypes, int32_t dim,
                       int32_t poly_deg, int32_t hv_k, double* M, double* R) {
    int np = orc_num_monomials(poly_deg);
    if (np < 0 || n <= 0) return -1;
    int m = n + np;
    int T = num_types(dertypes);
    Gauss rbf(eps, dim);
    std::memset(M, 0, sizeof(double) * (size_t)m * m);
    std::memset(R, 0, sizeof(double) * (size_t)m * T);
    // A(j,i) = phi(x_i - x_j)  (distanceMatrix, rbffd.cpp:947-964)
    for (int j = 0; j < n; j++) {
        const double* xj = xyz + 3 * (size_t)st[j];
        for (int i = 0; i < n; i++) {
            const double* xi = xyz + 3 * (size_t)st[i];
            M[(size_t)j * m + i] = rbf.eval(xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]);
        }
    }
    const double* x0 = xyz + 3 * (size_t)st[0];
    Mono mono[20];
    mono_list(poly_deg, mono);
    // P block: constant column of ones for the reference's np = 1 (rbffd.cpp:150-156)
    for (int p = 0; p < np; p++) {
        for (int i = 0; i < n; i++) {
            const double* xi = xyz + 3 * (size_t)st[i];
            double v = mono_eval(mono[p], xi[0] - x0[0], xi[1] - x0[1], xi[2] - x0[2]);
            M[(size_t)(n + p) * m + i] = v;
            M[(size_t)i * m + (n + p)] = v;
        }
    }
    // RHS columns in ascending DerType bit order (rbffd.cpp:106-125)
    int t = 0;
    for (int bit = 0; bit < 16; bit++) {
        uint32_t type = 1u << bit;
        if (!(dertypes & type)) continue;
        for (int j = 0; j < n; j++) R[(size_t)j * T + t] = rbf_rhs(type, rbf, x0, xyz + 3 * (size_t)st[j], hv_k);
        for (int p = 0; p < np; p++) R[(size_t)(n + p) * T + t] = mono_rhs(type, mono[p], x0, dim, hv_k);
        t++;
    }
    return m;
}

// Gaussian elimination with partial pivoting, outer-product form (Golub & Van Loan,
// Matrix Computations, Alg. 3.4.1), pivot = first row of maximal |a_ik| (LAPACK idamax rule),
// row interchanges applied to the RHS, then forward and back substitution. This is the
// algorithm of LAPACK dgesv, which arma::solve calls (rbffd.cpp:259).
int orc_lu_solve(int32_t m, int32_t nrhs, double* A, double* B) {
    for (int k = 0; k < m; k++) {
        int p = k;
        double amax = std::fabs(A[(size_t)k * m + k]);
        for (int i = k + 1; i < m; i++) {
            double v = std::fabs(A[(size_t)i * m + k]);
            if (v > amax) { amax = v; p = i; }
        }
        if (amax == 0.0) return k + 1;
        if (p != k) {
            for (int j = 0; j < m; j++) std::swap(A[(size_t)k * m + j], A[(size_t)p * m + j]);
            for (int j = 0; j < nrhs; j++) std::swap(B[(size_t)k * nrhs + j], B[(size_t)p * nrhs + j]);
        }
        double piv = A[(size_t)k * m + k];
        for (int i = k + 1; i < m; i++) {
            double l = A[(size_t)i * m + k] / piv;
            A[(size_t)i * m + k] = l;
            if (l == 0.0) continue;
            for (int j = k + 1; j < m; j++) A[(size_t)i * m + j] -= l * A[(size_t)k * m + j];
            for (int j = 0; j < nrhs; j++) B[(size_t)i * nrhs + j] -= l * B[(size_t)k * nrhs + j];
        }
    }
    for (int k = m - 1; k >= 0; k--) {
        for (int j = 0; j < nrhs; j++) {
            double s = B[(size_t)k * nrhs + j];
            for (int c = k + 1; c < m; c++) s -= A[(size_t)k * m + c] * B[(size_t)c * nrhs + j];
            B[(size_t)k * nrhs + j] = s / A[(size_t)k * m + k];
        }
    }
    return 0;
}

// RBFFD::computeAllWeightsForAllStencils -> computeAllWeightsForStencil_Direct
// (rbffd.cpp:90-103, 225-352): solve, scale the HV column by getHVScalar (295-297),
// keep the n node weights (the np multiplier entries are never applied, 669-688).
int orc_weights(int32_t n_st, const int32_t* st_off, const int32_t* st_idx, const double* xyz, const double* eps,
                uint32_t dertypes, int32_t dim, int32_t poly_deg, int32_t hv_k, double hv_scale, double* w,
                double* diag, int32_t n_threads) {
    int T = num_types(dertypes);
    size_t nnz = (size_t)st_off[n_st];
    int n_sing = 0;
#ifdef _OPENMP
    int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt) reduction(+ : n_sing)
#endif
    {
        std::vector<double> M, R, M0, R0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int s = 0; s < n_st; s++) {
            int n = st_off[s + 1] - st_off[s];
            int deg = orc_effective_degree(poly_deg, n);
            int m = n + orc_num_monomials(deg);
            M.resize((size_t)m * m);
            R.resize((size_t)m * T);
            orc_stencil_system(xyz, st_idx + st_off[s], n, eps[s], dertypes, dim, deg, hv_k, M.data(), R.data());
            if (diag) { M0 = M; R0 = R; }
            int info = orc_lu_solve(m, T, M.data(), R.data());
            if (info) n_sing++;
            int t = 0;
            for (int bit = 0; bit < 16; bit++) {
                uint32_t type = 1u << bit;
                if (!(dertypes & type)) continue;
                double scale = (type == 0x20) ? hv_scale : 1.0;
                for (int j = 0; j < n; j++)
                    w[(size_t)t * nnz + st_off[s] + j] = info ? NAN : R[(size_t)j * T + t] * scale;
                t++;
            }
            if (diag) {
                // normwise backward error per column, max over columns:
                // ||R - M W||_inf / (||M||_inf ||W||_inf + ||R||_inf)
                double mnorm = 0;
                for (int r = 0; r < m; r++) {
                    double rs = 0;
                    for (int c = 0; c < m; c++) rs += std::fabs(M0[(size_t)r * m + c]);
                    mnorm = std::max(mnorm, rs);
                }
                double be = 0;
                for (int tt = 0; tt < T; tt++) {
                    double rn = 0, wn = 0, bn = 0;
                    for (int r = 0; r < m; r++) {
                        double acc = R0[(size_t)r * T + tt];
                        for (int c = 0; c < m; c++) acc -= M0[(size_t)r * m + c] * R[(size_t)c * T + tt];
                        rn = std::max(rn, std::fabs(acc));
                        wn = std::max(wn, std::fabs(R[(size_t)r * T + tt]));
                        bn = std::max(bn, std::fabs(R0[(size_t)r * T + tt]));
                    }
                    be = std::max(be, rn / (mnorm * wn + bn));
                }
                double umax = 0, amax = 0;
                for (size_t q = 0; q < (size_t)m * m; q++) amax = std::max(amax, std::fabs(M0[q]));
                for (int r = 0; r < m; r++)
                    for (int c = r; c < m; c++) umax = std::max(umax, std::fabs(M[(size_t)r * m + c]));
                diag[2 * s] = info ? NAN : be;
                diag[2 * s + 1] = umax / amax;
            }
        }
    }
    return n_sing;
}

// RBFFD::applyWeightsForDeriv, CPU loop (rbffd.cpp:669-688): sequential sum in stencil order.
void orc_apply(int32_t n_rows, const int32_t* st_off, const int32_t* st_idx, const double* w, const double* u,
               double* out) {
    for (int i = 0; i < n_rows; i++) {
        double der = 0.0;
        for (int s = st_off[i]; s < st_off[i + 1]; s++) der += w[s] * u[st_idx[s]];
        out[i] = der;
    }
}

// One applyWeightsForDeriv per operator and field, over fixed-width stencils; rows in parallel.
void orc_spmm(int32_t n_rows, int32_t n, const int32_t* idx, int32_t n_ops, const double* w, int32_t n_fields,
              const double* u, double* out, int32_t n_threads) {
#ifdef _OPENMP
    int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (int i = 0; i < n_rows; i++) {
        for (int o = 0; o < n_ops; o++) {
            const double* wr = w + (size_t)o * n_rows * n + (size_t)i * n;
            const int32_t* ir = idx + (size_t)i * n;
            for (int f = 0; f < n_fields; f++) {
                double der = 0.0;
                for (int s = 0; s < n; s++) der += wr[s] * u[(size_t)ir[s] * n_fields + f];
                out[(size_t)o * n_rows * n_fields + (size_t)i * n_fields + f] = der;
            }
        }
    }
    (void)n_threads;
}

// Right-hand side of one stage for rows [0, n_rows): f[i][q] = sum_o c_o[i] (D_o y)[i][q], each
// dot product summed over the stencil in order j = 0..n-1 and the operators combined in order
// o = 0..n_ops-1 from 0.0 (the rounding of orc_spmm followed by that combination).
static void rhs_rows(int32_t n_rows, int32_t n, const int32_t* idx, int32_t n_ops, const double* w,
                     const double* const* coef, int32_t F, const double* y, double* f, int nt) {
    const size_t WN = (size_t)n_rows * n;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        std::vector<double> dot((size_t)F), acc((size_t)F);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int i = 0; i < n_rows; i++) {
            for (int q = 0; q < F; q++) acc[q] = 0.0;
            for (int o = 0; o < n_ops; o++) {
                const double* wr = w + (size_t)o * WN + (size_t)i * n;
                const int32_t* ir = idx + (size_t)i * n;
                for (int q = 0; q < F; q++) dot[q] = 0.0;
                for (int j = 0; j < n; j++) {
                    const double wj = wr[j];
                    const double* yr = y + (size_t)ir[j] * F;
                    for (int q = 0; q < F; q++) dot[q] += wj * yr[q];
                }
                const double c = (coef && coef[o]) ? coef[o][i] : 1.0;
                for (int q = 0; q < F; q++) acc[q] += (coef && coef[o]) ? c * dot[q] : dot[q];
            }
            for (int q = 0; q < F; q++) f[(size_t)i * F + q] = acc[q];
        }
    }
}

// TimeDependentPDE::advanceRungeKutta4 (time_dependent_pde.cpp:172-304): unscaled k's,
// s = u + 0.5*dt*k1, s = u + 0.5*dt*k2, s = u + dt*k3, u += dt/6*(k1 + 2k2 + 2k3 + k4).
// ws: 5 * n_rows * F doubles of caller-owned workspace (s, k1..k4), reused across steps.
void orc_rk4_step_ws(int32_t n_rows, int32_t n, const int32_t* idx, int32_t n_ops, const double* w,
                     const double* const* coef, int32_t F, double dt, double* u, double* ws, int32_t n_threads) {
    const long long NF = (long long)n_rows * F;
#ifdef _OPENMP
    const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#else
    const int nt = 1;
#endif
    double *s = ws, *k1 = ws + NF, *k2 = ws + 2 * NF, *k3 = ws + 3 * NF, *k4 = ws + 4 * NF;
    rhs_rows(n_rows, n, idx, n_ops, w, coef, F, u, k1, nt);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (long long i = 0; i < NF; i++) s[i] = u[i] + 0.5 * dt * (k1[i] + 0.);
    rhs_rows(n_rows, n, idx, n_ops, w, coef, F, s, k2, nt);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (long long i = 0; i < NF; i++) s[i] = u[i] + 0.5 * dt * (k2[i] + 0.);
    rhs_rows(n_rows, n, idx, n_ops, w, coef, F, s, k3, nt);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (long long i = 0; i < NF; i++) s[i] = u[i] + dt * (k3[i] + 0.);
    rhs_rows(n_rows, n, idx, n_ops, w, coef, F, s, k4, nt);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (long long i = 0; i < NF; i++) u[i] = u[i] + (dt / 6.) * (k1[i] + 2. * k2[i] + 2. * k3[i] + k4[i]);
}

void orc_rk4_step(int32_t n_rows, int32_t n, const int32_t* idx, int32_t n_ops, const double* w,
                  const double* const* coef, int32_t F, double dt, double* u, int32_t n_threads) {
    std::vector<double> ws((size_t)5 * n_rows * F);
    orc_rk4_step_ws(n_rows, n, idx, n_ops, w, coef, F, dt, u, ws.data(), n_threads);
}

void orc_cart2sph(double x, double y, double z, double* theta, double* phi, double* r) {
    Sph s = cart2sph(x, y, z);
    *theta = s.theta;
    *phi = s.phi;
    *r = s.r;
}

// Velocity coefficients of CosineBell::solve (cosine_bell.cpp:36-52), negated and divided by
// a so that f = c_lambda*Dlambda h + c_theta*Dtheta h + H h.
void orc_cosine_bell_coefs(int32_t n, const double* xyz, double a, double alpha, double t_rev, double* c_lambda,
                           double* c_theta) {
    double u0 = (2. * M_PI * a) / t_rev;  // cosine_bell.h:47
    for (int i = 0; i < n; i++) {
        Sph s = cart2sph(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        double lambda = s.theta, theta = s.phi;
        double vel_u = u0 * (std::cos(theta) * std::cos(alpha) + std::sin(theta) * std::cos(lambda) * std::sin(alpha));
        double vel_v = -u0 * (std::sin(lambda) * std::sin(alpha));
        c_lambda[i] = -(vel_u / a);
        c_theta[i] = -(vel_v / a);
    }
}

// ExactAdvection::operator() (tests/experiments/cosine_bell_vcl/exact_advection.h:37-54),
// h0 = 1, base 0, centre (3pi/2, 0), R = a/3; time-independent.
double orc_cosine_bell_exact(double x, double y, double z, double a) {
    const double lambda_c = (3. * M_PI) / 2., theta_c = 0., R = a / 3., h0 = 1., base_val = 0.;
    Sph s = cart2sph(x, y, z);
    double TI = s.phi, LI = s.theta;
    double r = a * std::acos(std::sin(theta_c) * std::sin(TI) + std::cos(theta_c) * std::cos(TI) * std::cos(LI - lambda_c));
    double h = (h0 / 2.) * (1. + std::cos(M_PI * (r / R))) + base_val;
    if (r >= R) h = base_val;
    return h;
}

// CosineBell::solve (cosine_bell.cpp:23-70) as written: three applies, then the pointwise
// velocity combination, then the HV filter.
void orc_cosine_bell_rhs(int32_t n_rows, const int32_t* st_off, const int32_t* st_idx, const double* w_lambda,
                         const double* w_theta, const double* w_hv, const double* xyz, double a, double alpha,
                         double t_rev, int32_t use_hv, const double* u, double* f) {
    std::vector<double> dl(n_rows), dt(n_rows), hv(n_rows);
    orc_apply(n_rows, st_off, st_idx, w_lambda, u, dl.data());
    orc_apply(n_rows, st_off, st_idx, w_theta, u, dt.data());
    if (use_hv) orc_apply(n_rows, st_off, st_idx, w_hv, u, hv.data());
    double u0 = (2. * M_PI * a) / t_rev;
    for (int i = 0; i < n_rows; i++) {
        Sph s = cart2sph(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        double lambda = s.theta, theta = s.phi;
        double vel_u = u0 * (std::cos(theta) * std::cos(alpha) + std::sin(theta) * std::cos(lambda) * std::sin(alpha));
        double vel_v = -u0 * (std::sin(lambda) * std::sin(alpha));
        f[i] = -((vel_u / (a)) * dl[i] + (vel_v / a) * dt[i]);
    }
    if (use_hv)
        for (int i = 0; i < n_rows; i++) f[i] += hv[i];
}

// PDE::calcSolNorms (pde.cpp:463-481) with l1/l2/linf of utils/norms.cpp:26-240.
void orc_sol_norms(int32_t n, const double* v, const double* ex, double* out) {
    double l1 = 0, l1d = 0, l2 = 0, l2d = 0, li = 0, lid = 0;
    for (int i = 0; i < n; i++) {
        double err = std::fabs(v[i] - ex[i]);
        l1 += err;
        l2 += err * err;
        li = (li < err) ? err : li;
        l1d += std::fabs(ex[i]);
        l2d += ex[i] * ex[i];
        lid = (lid < std::fabs(ex[i])) ? std::fabs(ex[i]) : lid;
    }
    l2 = std::sqrt(l2);
    l2d = std::sqrt(l2d);
    out[0] = l1;
    out[1] = (l1d > 1e-10) ? l1 / l1d : (l1 + 1.) / (l1d + 1.);
    out[2] = l2;
    out[3] = (l2d > 1e-10) ? l2 / l2d : (l2 + 1.) / (l2d + 1.);
    out[4] = li;
    out[5] = (lid > 1e-10) ? li / lid : (li + 1.) / (lid + 1.);
}

// This build's FP32 weight storage (no counterpart in the reference, whose FP32 mode is a
// plain cast, rbffd_cl.cpp:387-463): nearest rounding, then one pass over the slots in
// stencil order moving each weight by at most one FP32 ulp toward the remaining difference
// between the FP64 row sum and the sum of the FP32 values.
void orc_round_f32_rowsum(int32_t n_rows, int32_t width, const double* w, float* out) {
    for (int i = 0; i < n_rows; i++) {
        const double* src = w + (size_t)i * width;
        double s64 = 0.0, s32 = 0.0;
        for (int j = 0; j < width; j++) {
            s64 += src[j];
            s32 += (double)(float)src[j];
        }
        double r = s64 - s32;
        for (int j = 0; j < width; j++) {
            float v = (float)src[j];
            if (v != 0.0f && std::isfinite(r)) {
                float a = std::fabs(v);
                double ulp = (double)std::nextafter(a, INFINITY) - (double)a;
                double k = std::fmin(1.0, std::fmax(-1.0, std::rint(r / ulp)));
                if (k != 0.0 && std::isfinite(ulp)) {
                    float nv = (float)((double)v + k * ulp);
                    r -= (double)nv - (double)v;
                    v = nv;
                }
            }
            out[(size_t)i * width + j] = v;
        }
    }
}

}  // extern "C"
