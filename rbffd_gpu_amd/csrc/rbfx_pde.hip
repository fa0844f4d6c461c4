// Pointwise pieces of the PDEs advanced by the RK4 module.
//
// Cosine bell (src/pdes/hyperbolic/cosine_bell.{h,cpp}):
//   f = -((u_vel/a) Dlambda h + (v_vel/a) Dtheta h) + H h,
//   u_vel = u0 (cos(theta) cos(alpha) + sin(theta) cos(lambda) sin(alpha)),
//   v_vel = -u0 sin(lambda) sin(alpha), u0 = 2 pi a / T   (cosine_bell.cpp:36-52,
//   cosine_bell.h:47; lambda = atan2(y,x), theta = atan2(z, sqrt(x^2+y^2)), cart2sph.cpp:3-11).
// The RK4 stage kernel takes f = sum_o c_o (D_o h), so the velocity factors are precomputed
// once per node as c_lambda = -(u_vel/a), c_theta = -(v_vel/a) (like CosineBell_VCL's vel_u
// and vel_v vectors, cosine_bell_vcl.h:120-160); c_HV = 1.
// Initial condition: ExactAdvection (tests/experiments/cosine_bell_vcl/exact_advection.h:37-54).
#include "rbfx_internal.h"

namespace {

__global__ void cosine_bell_coefs_kernel(int n, const double* __restrict__ xyz, double a, double alpha, double u0,
                                         double* __restrict__ cl, double* __restrict__ ct) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    double lambda = atan2(y, x);
    double theta = atan2(z, sqrt(x * x + y * y));
    double vel_u = u0 * (cos(theta) * cos(alpha) + sin(theta) * cos(lambda) * sin(alpha));
    double vel_v = -u0 * (sin(lambda) * sin(alpha));
    cl[i] = -(vel_u / a);
    ct[i] = -(vel_v / a);
}

__global__ void cosine_bell_init_kernel(int n, const double* __restrict__ xyz, double a, double* __restrict__ h,
                                        int stride) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double lambda_c = (3. * M_PI) / 2., theta_c = 0., R = a / 3., h0 = 1., base = 0.;
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    double LI = atan2(y, x);
    double TI = atan2(z, sqrt(x * x + y * y));
    double r = a * acos(sin(theta_c) * sin(TI) + cos(theta_c) * cos(TI) * cos(LI - lambda_c));
    double v = (h0 / 2.) * (1. + cos(M_PI * (r / R))) + base;
    if (r >= R) v = base;
    h[(size_t)i * stride] = v;
}

}  // namespace

extern "C" {

int rbfx_cosine_bell_coefs(rbfx_ctx* ctx, int32_t n, const double* d_xyz, double a, double alpha, double t_rev,
                           double* d_c_lambda, double* d_c_theta) {
    if (!ctx) return RBFX_E_INVALID;
    if (n <= 0 || !d_xyz || !d_c_lambda || !d_c_theta || a == 0.0 || t_rev == 0.0)
        return rbfx_fail(ctx, RBFX_E_INVALID, "cosine_bell_coefs: bad arguments");
    double u0 = (2. * M_PI * a) / t_rev;
    cosine_bell_coefs_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, d_xyz, a, alpha, u0, d_c_lambda, d_c_theta);
    RBFX_LAUNCH_CHECK(ctx, "cosine_bell_coefs_kernel");
    return RBFX_OK;
}

int rbfx_cosine_bell_init(rbfx_ctx* ctx, int32_t n, const double* d_xyz, double a, double* d_h, int32_t stride) {
    if (!ctx) return RBFX_E_INVALID;
    if (n <= 0 || !d_xyz || !d_h || stride <= 0) return rbfx_fail(ctx, RBFX_E_INVALID, "cosine_bell_init: bad arguments");
    cosine_bell_init_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, d_xyz, a, d_h, stride);
    RBFX_LAUNCH_CHECK(ctx, "cosine_bell_init_kernel");
    return RBFX_OK;
}

}  // extern "C"
