/*
 * oracle.h - CPU restatement of bollig/rbffd_gpu's hot path, used ONLY as test
 * infrastructure: by tests/, by bench.py's cpu_baseline leg and by the smoke check, as the
 * checker. Nothing in the product (rbffd_gpu_amd/, include/) links or calls it.
 *
 * Each function names the reference file:line it restates. Layouts match the product's
 * C ABI so results compare directly:
 *   stencils  CSR: st_off[n_st+1], st_idx[nnz] (first entry of each stencil = centre)
 *   weights   w[t*nnz + st_off[s] + j], t = index of the t-th set DerType bit (ascending)
 *   fields    node-major u[node*n_fields + f]
 *
 * Pinning status of each function is recorded in the header comment of oracle.cpp.
 */
#ifndef RBFX_ORACLE_H
#define RBFX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gaussian RBF primitives (src/rbffd/rbfs/rbf_gaussian.h), evaluated at diff = (dx,dy,dz). */
double orc_rbf_eval(double eps, double dx, double dy, double dz);
double orc_rbf_xderiv(double eps, double dx, double dy, double dz);
double orc_rbf_lapl(double eps, int dim, double dx, double dy, double dz);
double orc_rbf_rderiv_over_r(double eps, double dx, double dy, double dz);
double orc_rbf_rderiv2(double eps, double dx, double dy, double dz);
double orc_rbf_hv(double eps, int dim, int k, double dx, double dy, double dz);
double orc_hv_pk(double eps, int dim, int k, double r2);

/* Number of polynomial terms for poly_deg in {-1,0,1,2,3}. */
int orc_num_monomials(int poly_deg);
/* Largest degree <= poly_deg with at most n monomials (used per stencil by orc_weights). */
int orc_effective_degree(int poly_deg, int n);

/* Builds the saddle-point system of one stencil (row-major m x m LHS, m x T RHS). */
int orc_stencil_system(const double* xyz, const int32_t* st, int32_t n, double eps, uint32_t dertypes,
                       int32_t dim, int32_t poly_deg, int32_t hv_k, double* M, double* R);

/* In-place partial-pivoting LU solve of M X = R (row-major, m x m and m x nrhs). Returns 0,
 * or k+1 when the k-th pivot is exactly zero. */
int orc_lu_solve(int32_t m, int32_t nrhs, double* M, double* R);

/* All weights (RBFFD::computeAllWeightsForAllStencils). diag may be NULL; else 2 doubles per
 * stencil {normwise backward error, LU pivot growth}. Returns number of singular stencils. */
int orc_weights(int32_t n_st, const int32_t* st_off, const int32_t* st_idx, const double* xyz,
                const double* eps, uint32_t dertypes, int32_t dim, int32_t poly_deg, int32_t hv_k,
                double hv_scale, double* w, double* diag, int32_t n_threads);

/* Derivative application (RBFFD::applyWeightsForDeriv CPU loop). */
void orc_apply(int32_t n_rows, const int32_t* st_off, const int32_t* st_idx, const double* w,
               const double* u, double* out);

/* Multi-operator, multi-field application with fixed width n (row-major idx/w per op):
 * out[o*n_rows*F + i*F + f] = sum_j w[o][i*n+j] * u[idx[i*n+j]*F + f]. */
void orc_spmm(int32_t n_rows, int32_t n, const int32_t* idx, int32_t n_ops, const double* w, int32_t n_fields,
              const double* u, double* out, int32_t n_threads);

/* One classic RK4 step (TimeDependentPDE::advanceRungeKutta4) of du/dt = f(u) with
 * f[i][q] = sum_o c_o[i] (D_o u)[i][q] (c_o NULL -> 1); u updated in place. */
void orc_rk4_step(int32_t n_rows, int32_t n, const int32_t* idx, int32_t n_ops, const double* w,
                  const double* const* coef, int32_t n_fields, double dt, double* u, int32_t n_threads);
/* The same step with caller-owned workspace ws of 5 * n_rows * n_fields doubles (s, k1..k4),
 * so that repeated steps allocate nothing; rows and vector updates run on n_threads. */
void orc_rk4_step_ws(int32_t n_rows, int32_t n, const int32_t* idx, int32_t n_ops, const double* w,
                     const double* const* coef, int32_t n_fields, double dt, double* u, double* ws,
                     int32_t n_threads);

/* Cosine bell (src/pdes/hyperbolic/cosine_bell.{h,cpp}, exact_advection.h). */
void orc_cart2sph(double x, double y, double z, double* theta, double* phi, double* r);
void orc_cosine_bell_coefs(int32_t n, const double* xyz, double a, double alpha, double t_rev,
                           double* c_lambda, double* c_theta);
double orc_cosine_bell_exact(double x, double y, double z, double a);
void orc_cosine_bell_rhs(int32_t n_rows, const int32_t* st_off, const int32_t* st_idx, const double* w_lambda,
                         const double* w_theta, const double* w_hv, const double* xyz, double a, double alpha,
                         double t_rev, int32_t use_hv, const double* u, double* f);

/* PDE::calcSolNorms (src/pdes/pde.cpp:463-533): out = {l1abs, l1rel, l2abs, l2rel, liabs, lirel}. */
void orc_sol_norms(int32_t n, const double* approx, const double* exact, double* out6);

/* FP32 weight storage keeping the FP64 row sums (this build's rounding; row-major in/out). */
void orc_round_f32_rowsum(int32_t n_rows, int32_t width, const double* w, float* out);

/* Steady Stokes (oracle_stokes.cpp). Matrices are CSR with int64 row offsets and ascending
 * columns. w holds the LSFC, XSFC, YSFC, ZSFC weights of fixed-width stencils (4 x N*n,
 * row-major, idx[i*n + j]). */
int64_t orc_stokes_nnz(int32_t N, int32_t n);
void orc_stokes_assemble(int32_t N, int32_t n, const int32_t* idx, const double* w, double eta, int64_t* ip,
                         int32_t* ind, double* val);
void orc_csr_matvec(int32_t nrows, const int64_t* ip, const int32_t* ind, const double* val, const double* x,
                    double* y);
/* ILU(0) restricted to rows/columns [row_start, row_end); returns the factor's nnz
 * (lip/lind/lval sized like the input). */
int64_t orc_ilu0(int32_t nrows, const int64_t* ip, const int32_t* ind, const double* val, int32_t row_start,
                 int32_t row_end, int64_t* lip, int32_t* lind, double* lval);
void orc_ilu0_apply(int32_t nrows, const int64_t* lip, const int32_t* lind, const double* lval, double* v);
/* GMRES(krylov) with left ILU(0) preconditioning (lip NULL: none). Returns 0 when converged. */
int orc_gmres(int32_t nrows, const int64_t* ip, const int32_t* ind, const double* val, const int64_t* lip,
              const int32_t* lind, const double* lval, const double* b, double tol, int32_t max_iter,
              int32_t krylov, double* x, int32_t* iters, double* error);

#ifdef __cplusplus
}
#endif
#endif
