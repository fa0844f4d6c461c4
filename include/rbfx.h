/*
 * rbfx.h - C ABI of the MI355X (gfx950) RBF-FD hot path.
 *
 * Plain C: pointers and sizes only. Every function returns RBFX_OK (0) or a negative
 * RBFX_E_* code and never exits the process; rbfx_last_error() gives the message.
 * A context owns one HIP stream and all device buffers it creates; it is single-threaded
 * (like one rank of the reference) and asynchronous on its stream.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * bollig/rbffd_gpu tree).
 */
#ifndef RBFX_H
#define RBFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------- */
#define RBFX_OK 0
#define RBFX_E_INVALID -1      /* bad argument (null, size, mask, unsupported option)   */
#define RBFX_E_HIP -2          /* a HIP runtime call failed                              */
#define RBFX_E_NOMEM -3        /* device or host allocation failed                       */
#define RBFX_E_NODEVICE -4     /* no usable GPU                                          */
#define RBFX_E_SINGULAR -5     /* at least one stencil system had a zero pivot           */
#define RBFX_E_STATE -6        /* call out of order (e.g. step before state was set)     */
#define RBFX_E_TOO_LARGE -7    /* size beyond a documented limit                         */

/* ---- derivative types: bit values of RBFFD::DerType (src/rbffd/rbffd.h:43-60) -------- */
#define RBFX_X 0x1u
#define RBFX_Y 0x2u
#define RBFX_Z 0x4u
#define RBFX_LAPL 0x8u
#define RBFX_R 0x10u
#define RBFX_HV 0x20u
#define RBFX_LAMBDA 0x40u
#define RBFX_THETA 0x80u
#define RBFX_LSFC 0x100u
#define RBFX_XSFC 0x200u
#define RBFX_YSFC 0x400u
#define RBFX_ZSFC 0x800u
#define RBFX_INTERP 0x8000u
/* ALT_XSFC/ALT_YSFC/ALT_ZSFC (0x1000-0x4000) are linear combinations of X,Y,Z weights
 * (src/rbffd/rbffd.cpp:337-347); they are formed on the host side, not solved for. */

/* ---- precision selectors ------------------------------------------------------------- */
#define RBFX_F64 0
#define RBFX_F32 1
/* Weight storage only (rbfx_ell_create): FP32 with plain round-to-nearest, the reference's
 * (float) cast (src/rbffd/rbffd_cl.cpp:387-463). RBFX_F32 weight storage instead rounds each
 * row so that the FP32 weights keep the FP64 row sum (each weight within 1.5 ulp). */
#define RBFX_F32_RN 2

/* weight-solve modes */
#define RBFX_SOLVE_F64_LU 0          /* FP64 partial-pivot LU in LDS                       */
#define RBFX_SOLVE_F32_LU_F64_IR 1   /* FP32 LU + FP64-residual refinement, FP64 fallback  */

/* poly_deg flag of rbfx_weights: sphere-reduced monomial basis for nodes exactly on the unit
 * sphere (degree 2: 9 functions, degree 3: 16; the full basis is singular there). */
#define RBFX_POLY_SPHERE 16

/* per-stencil diagnostic path flags (diag_out[4*s+3]) */
#define RBFX_PATH_F64 0
#define RBFX_PATH_F32_IR 1
#define RBFX_PATH_F32_FALLBACK_F64 2
#define RBFX_PATH_SINGULAR 3

typedef struct rbfx_ctx rbfx_ctx;
typedef struct rbfx_ell rbfx_ell;
typedef struct rbfx_rk4 rbfx_rk4;

/* ---- context ------------------------------------------------------------------------- */
/* Replaces the OpenCL context/queue set-up of CLBaseClass (src/utils/opencl/cl_base_class.h:14-48). */
int rbfx_create(int device_ordinal, rbfx_ctx** out);
void rbfx_destroy(rbfx_ctx* ctx);
/* Message of the last failure on this context (or of rbfx_create when ctx is NULL). */
const char* rbfx_last_error(const rbfx_ctx* ctx);
/* Blocks until all work queued on the context's stream is done (queue.finish()). */
int rbfx_sync(rbfx_ctx* ctx);
/* Device name and peak figures as reported by the HIP runtime. */
int rbfx_device_info(rbfx_ctx* ctx, char* name, int32_t name_len, int32_t* n_cu, int32_t* clock_khz);

/* ---- device memory helpers (the ctx's stream; host pointers are pageable) ------------ */
int rbfx_dev_alloc(rbfx_ctx* ctx, uint64_t bytes, void** d_ptr);
int rbfx_dev_free(rbfx_ctx* ctx, void* d_ptr);
int rbfx_h2d(rbfx_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes);
int rbfx_d2h(rbfx_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes);
int rbfx_d2d(rbfx_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes);

/* Plain device-to-device streaming copy kernel (16-byte aligned pointers and size): the
 * measured HBM reference for the roofline figures. No counterpart in the reference. */
int rbfx_stream_copy(rbfx_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes);

/* Stream-ordered timer: records a HIP event pair on the ctx stream around the work queued
 * between start and stop; stop waits for completion and returns elapsed milliseconds. */
int rbfx_timer_start(rbfx_ctx* ctx);
int rbfx_timer_stop(rbfx_ctx* ctx, double* ms);

/* ---- host-side stencil generation ---------------------------------------------------- */
/* Grid::generateStencilsKDTree (src/grids/grid_interface.cpp:688-725) with the reference's
 * kd-tree (external/kdtree/KDTree.h): out_idx[i*k + j] = j-th nearest node of node i
 * (j = 0 is node i itself), bit-identical to the reference, ties included.
 * Host memory, CPU work; n_threads <= 0 uses all OpenMP threads. */
int rbfx_knn_kdtree(int32_t n_nodes, const double* xyz, int32_t k, int32_t* out_idx, int32_t n_threads);
/* The same stencils from a GPU search (replaces the query loop of Grid::generateStencilsKDTree,
 * src/grids/grid_interface.cpp:688-725; host arrays in and out). A query whose k+1 nearest
 * distances are all distinct has one answer, the k nearest in ascending distance, which the
 * GPU computes exactly over a uniform grid of cells with the kd-tree's distance arithmetic.
 * Queries with a distance tie among their k+1 nearest (the kd-tree's heap order decides
 * those) are recomputed on the host by the kd-tree; n_host_rows (may be NULL) receives how
 * many. The result is bit-identical to rbfx_knn_kdtree. k <= n_nodes. */
int rbfx_knn_gpu(rbfx_ctx* ctx, int32_t n_nodes, const double* xyz, int32_t k, int32_t* out_idx, int32_t* n_host_rows);
/* rbfx_knn_gpu with the lists written to device memory instead (d_out_idx: n_nodes * k int32
 * on the context's device; xyz stays a host array), so that rbfx_weights and rbfx_ell_create
 * can take them without a round trip through host memory. Tie rows are recomputed on the
 * host and uploaded. No counterpart in the reference (its stencils live on the host). */
int rbfx_knn_gpu_device(rbfx_ctx* ctx, int32_t n_nodes, const double* xyz, int32_t k, int32_t* d_out_idx,
                        int32_t* n_host_rows);
/* Seeded node set on the unit sphere: golden-angle (Fibonacci) spiral plus an independent
 * uniform +-jitter per coordinate (SplitMix64 counter stream), not re-projected.
 * Stands in for GridReader::readNodeList's input files (src/grids/grid_reader.cpp:36-100). */
int rbfx_nodes_fibonacci(int32_t n_nodes, double jitter, uint64_t seed, double* xyz);
/* Icosahedral node set of the given subdivision level, exactly on the unit sphere:
 * N = 10*4^level + 2 (level 7 = 163842, the quasi-uniform cosine-bell grid size).
 * With xyz == NULL only *out_n is set. */
int rbfx_nodes_icosahedral(int32_t level, double* xyz, int32_t* out_n);
/* Host: order of the nodes along a 3-D Morton (Z-order) curve over their bounding box
 * (21 bits per axis, ties by index): out_order[p] = node at position p. For
 * rbfx_ell_set_row_order. No counterpart in the reference. */
int rbfx_order_morton(int32_t n_nodes, const double* xyz, int32_t* out_order);

/* Real spherical harmonic Y_lm at the directions of n points (host arrays; out[i] for point i):
 * orthonormal on the unit sphere, Condon-Shortley phase included, cos(m phi) for m > 0 and
 * sin(|m| phi) for m < 0 - the normalisation and signs of the closed forms in
 * src/utils/spherical_harmonics.h:39-170 (Y_3^2, Y_10^5, Y_20^20), for any 0 <= |m| <= l.
 * Field generator of the multi-field workloads (seeded combinations of degree <= 20). */
int rbfx_sph_harm(int32_t l, int32_t m, int32_t n, const double* xyz, double* out);

/* ---- batched weight computation ------------------------------------------------------ */
/* RBFFD::computeAllWeightsForAllStencils -> computeAllWeightsForStencil_Direct
 * (src/rbffd/rbffd.cpp:90-103, 225-352) with getStencilLHS/distanceMatrix (129-191, 888-969)
 * and getStencilMultiRHS/getStencilRHS (106-125, 356-537), for all stencils in one call.
 *
 * Stencils are CSR-shaped and may have uneven sizes: stencil s uses nodes
 * d_st_idx[d_st_off[s] .. d_st_off[s+1]), first entry = centre. All pointers are device
 * pointers: d_xyz N*3 float64, d_eps one shape parameter per stencil.
 * dertypes: RBFX_* mask; one output column per set bit in ascending bit order.
 * dim: dimension of LAPL and of the HV recurrence (RBFFD(..., dim_num, ...)).
 * poly_deg: -1 = no polynomial, 0 = the reference's single constant monomial (ONE_MONOMIAL,
 *   rbffd.cpp:1), 1..3 = full monomial basis in centred coordinates (x - x_centre). A stencil
 *   with fewer nodes than monomials uses the largest degree that fits (n = 17 with degree 3
 *   uses degree 2), since [A P; P' 0] is singular otherwise.
 * hv_k, hv_scale: Delta^k order and the column scale -gamma*N^-k (RBFFD::getHVScalar,
 *   rbffd.h:227-238); hv_scale multiplies the HV column only.
 * Output (device): d_w[t * nnz + d_st_off[s] + j] for type column t, node j < n_s
 *   (the polynomial multipliers are not stored). d_diag (optional, may be NULL):
 *   4 doubles per stencil {backward error ||R - M W||_max/(||M||_max ||W||_max), max |pivot|
 *   ratio, refinement iterations, path flag}.
 * n_singular (optional host pointer) receives the count of stencils with a zero pivot;
 * their weights are left as NaN and the call returns RBFX_E_SINGULAR. */
int rbfx_weights(rbfx_ctx* ctx, int32_t n_stencils, const int32_t* d_st_off, const int32_t* d_st_idx,
                 int32_t n_nodes, const double* d_xyz, const double* d_eps, uint32_t dertypes, int32_t dim,
                 int32_t poly_deg, int32_t hv_k, double hv_scale, int32_t mode, double* d_w, double* d_diag,
                 int32_t* n_singular);

/* rbfx_weights that allocates no device memory. No counterpart in the reference: it is for
 * callers that cannot allocate (code under stream capture, serving loops, harnesses that
 * forbid it). Same arguments, same kernels, same results. The solver's scratch (stencil order,
 * per-stencil flags, size-class table, work counter: about 8 bytes per stencil) lives in a
 * file-scope device array sized for RBFX_WEIGHTS_NOALLOC_MAX_STENCILS stencils, which each
 * device holds until the process ends; n_stencils above it is RBFX_E_TOO_LARGE. Every call on
 * a device shares that device's array, so calls that overlap on one device (several contexts
 * or threads) race and must be serialised by the caller. Host-side staging (the size sort) still
 * uses host memory. */
#define RBFX_WEIGHTS_NOALLOC_MAX_STENCILS 262144
int rbfx_weights_noalloc(rbfx_ctx* ctx, int32_t n_stencils, const int32_t* d_st_off, const int32_t* d_st_idx,
                         int32_t n_nodes, const double* d_xyz, const double* d_eps, uint32_t dertypes, int32_t dim,
                         int32_t poly_deg, int32_t hv_k, double hv_scale, int32_t mode, double* d_w, double* d_diag,
                         int32_t* n_singular);

/* ---- ELL operators (fixed stencil width) --------------------------------------------- */
/* Builds the device layout of n_ops operators that share one index array, the analogue of
 * RBFFD_CL::updateStencilsOnGPU/updateWeightsOnGPU (src/rbffd/rbffd_cl.cpp:215-257, 306-463)
 * and RBFFD_VCL's ell_matrix (src/rbffd/rbffd_vcl.cpp:224-299).
 * Inputs (device): d_idx n_rows*width int32 row-major (row i, slot j; pad slots = centre),
 * d_w[o] n_rows*width float64 row-major (pad weight 0); storage RBFX_F64, RBFX_F32
 * (row-sum-preserving rounding) or RBFX_F32_RN (nearest); both FP32 modes report RBFX_F32.
 * dertype_bits[o] (host) labels operator o. */
int rbfx_ell_create(rbfx_ctx* ctx, int32_t n_rows, int32_t n_cols, int32_t width, const int32_t* d_idx,
                    int32_t n_ops, const uint32_t* dertype_bits, const double* const* d_w, int32_t storage,
                    rbfx_ell** out);
void rbfx_ell_destroy(rbfx_ell* ell);
/* Processing order of the rows (no counterpart in the reference, whose rows run in node order):
 * d_order (device, n_rows int32, a permutation) lists the row handled at each position, and
 * the operator set is re-laid out in that order. Results are unchanged bit for bit (every row
 * is still computed the same way and written at its own index); what changes is which rows
 * run side by side. With rows ordered along a space-filling curve (rbfx_order_morton), the
 * rows of one wave share most of their stencil nodes, so the neighbour gathers hit in the
 * CU's L1 instead of each going to L2. RK4 steppers on this set recapture their graph. */
int rbfx_ell_set_row_order(rbfx_ell* ell, const int32_t* d_order);
/* LDS-staged gathers (no counterpart in the reference): cuts the positions greedily into
 * tiles of at most tile_rows (32, 64 or 128) consecutive rows with at most max_nodes distinct
 * stencil nodes each (0 = no limit; at least the stencil width), and records per tile its
 * node list and 16-bit local indices into it. The RK4 stage kernel's shape share = 4
 * (rbfx_tune_set) then copies a tile's neighbour records into LDS once and gathers from
 * there; max_nodes bounds that LDS (max_nodes * n_fields * field bytes per block). Build
 * after rbfx_ell_set_row_order (re-ordering drops the tables); with rows in a locality order
 * a 64-row tile has a few hundred distinct neighbours. rbfx_ell_local_info reports the tile
 * size (0 = not built), the largest tile list and the number of tiles. */
int rbfx_ell_build_local(rbfx_ell* ell, int32_t tile_rows, int32_t max_nodes);
int rbfx_ell_local_info(const rbfx_ell* ell, int32_t* tile_rows, int32_t* k_max, int32_t* n_tiles);
int rbfx_ell_info(const rbfx_ell* ell, int32_t* n_rows, int32_t* width, int32_t* n_ops, int32_t* storage);

/* Derivative application RBFFD::applyWeightsForDeriv (src/rbffd/rbffd.cpp:669-688,
 * rbffd_cl.cpp:526-694): out[i] = sum_j W_op[i][j] * u[idx[i][j]] for operator index op.
 * d_u: n_cols values, d_out: n_rows values, both of type field_type (RBFX_F64/F32);
 * accumulation in accum (RBFX_F64/F32). */
int rbfx_spmv(rbfx_ell* ell, int32_t op, const void* d_u, void* d_out, int32_t field_type, int32_t accum);

/* Fused multi-operator, multi-field application:
 * out[o][i][f] = sum_j W_o[i][j] * u[idx[i][j]][f] for every operator o of the set and
 * field f < n_fields; u is node-major (n_cols x n_fields), out is n_ops x n_rows x n_fields. */
int rbfx_spmm(rbfx_ell* ell, int32_t n_fields, const void* d_u, void* d_out, int32_t field_type, int32_t accum);

/* ---- RK4 time stepping ---------------------------------------------------------------- */
/* TimeDependentPDE::advanceRungeKutta4 (src/pdes/time_dependent_pde.cpp:172-304), the OpenCL
 * evaluateRK4_* / advanceRK4_classic kernels (src/cl_kernels/rk4_thread_per_stencil.cl:13-138)
 * and TimeDependentPDE_VCL::advanceRK4 (src/pdes/time_dependent_pde_vcl.cpp:246-393), with
 * the PDE right-hand side as a per-node linear combination of the operators:
 *     f[i][q] = sum_o c_o[i] * (D_o u)[i][q]
 * where c_o is a device array of n_rows float64 or NULL for the constant 1. The cosine bell
 * (src/pdes/hyperbolic/cosine_bell.cpp:23-70) is c_LAMBDA = -u_vel/a, c_THETA = -v_vel/a,
 * c_HV = 1. Each step runs 4 fused stage kernels (gather-SpMM + right-hand side + stage
 * update) captured once into a hipGraph. State u is float64, node-major n_rows x n_fields. */
typedef struct rbfx_rk4_desc {
    rbfx_ell* ell;             /* operators (n_rows == n_cols)                              */
    int32_t n_fields;          /* fields advanced together                                  */
    int32_t field_storage;     /* RBFX_F64: stage fields in FP64; RBFX_F32: stored in FP32   */
    int32_t accum;             /* RBFX_F64 or RBFX_F32 accumulation of the gathers          */
    const double* d_coef[16];  /* per-operator coefficient arrays (NULL = 1), n_ops used    */
    int32_t use_graph;         /* 1: capture a step into a hipGraph and replay it           */
} rbfx_rk4_desc;

int rbfx_rk4_create(rbfx_ctx* ctx, const rbfx_rk4_desc* desc, rbfx_rk4** out);
void rbfx_rk4_destroy(rbfx_rk4* rk);
/* TimeDependentPDE_HIP::fillInitialConditions / syncCPUtoGPU analogues (device pointers). */
int rbfx_rk4_set_state(rbfx_rk4* rk, const double* d_u);
int rbfx_rk4_get_state(rbfx_rk4* rk, double* d_u);
/* Advances n_steps steps of size dt (TimeDependentPDE::advance(RK4, dt) repeated). */
int rbfx_rk4_step(rbfx_rk4* rk, double dt, int32_t n_steps);
/* One stage kernel launch (stage 1..4) for profiling the dominant kernel in isolation.
 * It updates the internal state like the stage would inside a step. */
int rbfx_rk4_stage(rbfx_rk4* rk, int32_t stage, double dt);

/* ---- launch-shape tuning (fused kernels; kind 0 = SpMM, 1 = RK4 stage) ----------------- */
/* Fixes fields-per-thread (1,2,4,8), block size (64..256, multiple of 64) and gather batch
 * (4 or 8 stencil slots per memory round trip; 16 for SpMV/SpMM with one field per thread)
 * for this operator set's row count, width,
 * operator count, field count and precisions. rbfx_rk4_step picks the shape up when it
 * (re)captures its graph. No counterpart in the reference (its launch shape is fixed). */
/* share = 1 (RK4 stage kernel only): the G = n_fields/fpt lanes of a row (G = 2 or 4, fpt > 1,
 * batch divisible by G) split the index and weight loads of each slot batch and broadcast
 * them inside the group, instead of every lane loading them. share = 2 or 3: the same, with
 * those loads issued 2 or 3 batches ahead and the next batch's gathers issued before the
 * current batch is consumed (3-4 operators; fpt 4 with 4 lanes or fpt 8 with 2 lanes, and
 * for depth 2 also fpt 4 with 2 lanes or fpt 2 with 4 lanes). share = 4: LDS-staged gathers
 * (needs rbfx_ell_build_local; the block is the tile). share = 5: share 4 with the FP32
 * multiply-adds on the matrix cores (FP32 weights, fields and accumulation, fpt 4 with 4
 * lanes, i.e. 16 fields); bit-identical to share 4. */
int rbfx_tune_set(rbfx_ell* ell, int32_t kind, int32_t n_fields, int32_t field_type, int32_t accum, int32_t fpt,
                  int32_t block, int32_t batch, int32_t share);
int rbfx_tune_get(rbfx_ell* ell, int32_t kind, int32_t n_fields, int32_t field_type, int32_t accum, int32_t* fpt,
                  int32_t* block, int32_t* batch, int32_t* share);
int rbfx_tune_clear(void);

/* ---- steady Stokes on the sphere ------------------------------------------------------- */
/* StokesSteadyPDE (src/pdes/elliptic/stokes_steady_pde.cpp:61-230): the coupled system in the
 * reference's block layout, unknowns u (0..N), v (N..2N), w (2N..3N), p (3N..4N) and the four
 * constraint multipliers (4N..4N+3):
 *   -eta L u + D_x p + lambda_u = f_u, ..., D_x u + D_y v + D_z w + lambda_p = f_p,
 *   sum u = g_u, ..., sum p = g_p,
 * with L, D_x, D_y, D_z the LSFC, XSFC, YSFC, ZSFC weights. ops64 holds those four operators
 * (types in that order) in FP64 storage on square fixed-width stencils that contain their
 * centre; ops32 (optional) the same weights in FP32 storage, for the FP32 inner solves.
 * Creating the handle factors the preconditioner of ilu0_precond with ilu0_tag(0, 3N)
 * (src/precond/ilu0.hpp:157-260, tests/util/viennacl_gmres_poisson/main_stokes.cpp:83): ILU(0)
 * of the velocity block, identity rows for the pressure and constraint rows. The operator sets
 * must outlive the handle. */
typedef struct rbfx_stokes rbfx_stokes;
int rbfx_stokes_create(rbfx_ctx* ctx, rbfx_ell* ops64, rbfx_ell* ops32, double eta, rbfx_stokes** out);
void rbfx_stokes_destroy(rbfx_stokes* s);
/* Unknowns (4N + 4) and the number of steps of the two triangular sweeps: their dependency
 * levels, each split into steps of at most 32 rows (16 for stencils above 33 nodes). */
int rbfx_stokes_info(const rbfx_stokes* s, int32_t* n_unknowns, int32_t* lower_levels, int32_t* upper_levels);
/* d_y = A d_x (device, 4N + 4 doubles each): the reference's prod(LHS, x). */
int rbfx_stokes_apply(rbfx_stokes* s, const double* d_x, double* d_y);
/* The ILU(0) factor of -eta L (N rows of the stencil width, columns ascending per row, L
 * multipliers below the diagonal, U from it): the velocity rows of the reference's LU. */
int rbfx_stokes_factor(rbfx_stokes* s, double* d_f);
/* d_v = M^-1 d_v (ilu0_lu_substitute, src/precond/ilu0.hpp:273-327). */
int rbfx_stokes_precond(rbfx_stokes* s, double* d_v);
/* viennacl::linalg::solve(LHS, RHS, gmres_tag(tol, max_iter, krylov), ilu0_precond) (ViennaCL
 * gmres.hpp, the reference's external/opencl_libraries/include/viennacl/linalg/gmres.hpp:166-336):
 * restarted GMRES(krylov) from x = 0 with left preconditioning, stopping when the estimate of
 * ||M^-1 (b - A x)|| / ||b|| falls below tol or after the restarts ViennaCL allows
 * (max_iter / krylov cycles, one fewer when krylov divides max_iter). mode RBFX_F64: all in
 * FP64. mode RBFX_F32 (needs ops32): FP64 refinement, r = b - A x and the stopping test in
 * FP64, each correction from one FP32 GMRES(krylov) cycle (FP32 operator, factor and basis)
 * that reduces its preconditioned residual 1e4-fold, at most max_iter inner iterations in all.
 * iters receives the Arnoldi steps taken, error the final relative preconditioned residual
 * (the estimate in FP64 mode, recomputed in FP64 in FP32 mode). Not converging is not an
 * error: compare error with tol. */
int rbfx_stokes_gmres(rbfx_stokes* s, const double* d_b, double* d_x, double tol, int32_t max_iter, int32_t krylov,
                      int32_t mode, int32_t* iters, double* error);

/* ---- cosine bell pointwise terms (src/pdes/hyperbolic/cosine_bell.cpp:23-70) ---------- */
/* c_lambda[i] = -(u_vel/a), c_theta[i] = -(v_vel/a) at node i (device arrays), for use as the
 * LAMBDA / THETA coefficients of rbfx_rk4_desc (the HV coefficient is 1). */
int rbfx_cosine_bell_coefs(rbfx_ctx* ctx, int32_t n, const double* d_xyz, double a, double alpha, double t_rev,
                           double* d_c_lambda, double* d_c_theta);
/* Bell height (ExactAdvection, tests/experiments/cosine_bell_vcl/exact_advection.h:37-54) at
 * each node, written to d_h[i*stride]. */
int rbfx_cosine_bell_init(rbfx_ctx* ctx, int32_t n, const double* d_xyz, double a, double* d_h, int32_t stride);

#ifdef __cplusplus
}
#endif

#endif /* RBFX_H */
