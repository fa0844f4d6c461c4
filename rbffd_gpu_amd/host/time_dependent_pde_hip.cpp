#include "time_dependent_pde_hip.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>

namespace rbfx {

namespace {
[[noreturn]] void die(rbfx_ctx* ctx, const char* what, int rc) {
    std::cout << "[TimeDependentPDE_HIP] ERROR in " << what << " (" << rc << "): " << rbfx_last_error(ctx) << std::endl;
    std::exit(EXIT_FAILURE);
}
#define RBFX_CALL(ctx, call)                      \
    do {                                          \
        int rc_ = (call);                         \
        if (rc_ != RBFX_OK) die(ctx, #call, rc_); \
    } while (0)
}  // namespace

double ExactAdvection::operator()(double x, double y, double z, double) {
    const double lambda_c = (3. * M_PI) / 2., theta_c = 0., h0 = 1., base_val = 0.;
    const double R = a / 3.;
    double TI = std::atan2(z, std::sqrt(x * x + y * y));
    double LI = std::atan2(y, x);
    double r = a * std::acos(std::sin(theta_c) * std::sin(TI) + std::cos(theta_c) * std::cos(TI) * std::cos(LI - lambda_c));
    double h = (h0 / 2.) * (1. + std::cos(M_PI * (r / R))) + base_val;
    if (r >= R) h = base_val;
    return h;
}

TimeDependentPDE_HIP::TimeDependentPDE_HIP(Domain* grid, RBFFD_HIP* der, Communicator*, int gpu_type_, bool weightsComputed_)
    : grid_ref(*grid), der_ref(*der), gpu_type(gpu_type_), weightsComputed(weightsComputed_) {
    U_G.assign(grid_ref.getNodeListSize(), 0.0);
}

TimeDependentPDE_HIP::~TimeDependentPDE_HIP() {
    rbfx_rk4_destroy(rk);
    rbfx_dev_free(der_ref.getContext(), d_state);
}

void TimeDependentPDE_HIP::assemble() {
    if (!weightsComputed) {
        der_ref.computeAllWeightsForAllStencils();
        weightsComputed = true;
    }
}

void TimeDependentPDE_HIP::fillInitialConditions(ExactSolution* exact) {
    exact_ptr = exact;
    std::vector<NodeType>& nodes = grid_ref.getNodeList();
    for (size_t i = 0; i < nodes.size(); i++) U_G[i] = exact ? exact->at(nodes[i], cur_time) : 0.0;
    assemble();
    ensureStepper();
    rbfx_ctx* ctx = der_ref.getContext();
    RBFX_CALL(ctx, rbfx_h2d(ctx, d_state, U_G.data(), U_G.size() * sizeof(double)));
    RBFX_CALL(ctx, rbfx_rk4_set_state(rk, d_state));
    state_on_device = true;
}

void TimeDependentPDE_HIP::ensureStepper() {
    if (rk) return;
    rbfx_ctx* ctx = der_ref.getContext();
    der_ref.setGPUPrecision(field_storage, accum);
    rbfx_rk4_desc d = {};
    d.ell = der_ref.getGPUOperatorSet(rhsOperators());
    d.n_fields = 1;
    d.field_storage = field_storage;
    d.accum = accum;
    std::vector<const double*> c;
    rhsCoefficients(c);
    for (size_t o = 0; o < c.size() && o < 16; o++) d.d_coef[o] = c[o];
    d.use_graph = 1;
    RBFX_CALL(ctx, rbfx_rk4_create(ctx, &d, &rk));
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, U_G.size() * sizeof(double), (void**)&d_state));
}

void TimeDependentPDE_HIP::advance(TimeScheme which, double delta_t) {
    if (which != RK4) {
        std::cout << "[TimeDependentPDE_HIP] Only the RK4 scheme is provided on the GPU. Bailing...\n";
        std::exit(EXIT_FAILURE);
    }
    advanceSteps(delta_t, 1);
}

void TimeDependentPDE_HIP::advanceSteps(double delta_t, int n_steps) {
    if (!state_on_device) fillInitialConditions(exact_ptr);
    RBFX_CALL(der_ref.getContext(), rbfx_rk4_step(rk, delta_t, n_steps));
    cur_time += delta_t * n_steps;
}

void TimeDependentPDE_HIP::syncCPUtoGPU() {
    if (!state_on_device) return;
    rbfx_ctx* ctx = der_ref.getContext();
    RBFX_CALL(ctx, rbfx_rk4_get_state(rk, d_state));
    RBFX_CALL(ctx, rbfx_d2h(ctx, U_G.data(), d_state, U_G.size() * sizeof(double)));
}

std::vector<double> TimeDependentPDE_HIP::checkGlobalError(ExactSolution* exact, double rel_err_max, bool ignore) {
    syncCPUtoGPU();
    std::vector<NodeType>& nodes = grid_ref.getNodeList();
    double l1 = 0, l1d = 0, l2 = 0, l2d = 0, li = 0, lid = 0;
    for (size_t i = 0; i < nodes.size(); i++) {
        double ex = exact->at(nodes[i], cur_time);
        double err = std::fabs(U_G[i] - ex);
        l1 += err;
        l2 += err * err;
        li = li < err ? err : li;
        l1d += std::fabs(ex);
        l2d += ex * ex;
        lid = lid < std::fabs(ex) ? std::fabs(ex) : lid;
    }
    l2 = std::sqrt(l2);
    l2d = std::sqrt(l2d);
    double l1r = l1d > 1e-10 ? l1 / l1d : (l1 + 1.) / (l1d + 1.);
    double l2r = l2d > 1e-10 ? l2 / l2d : (l2 + 1.) / (l2d + 1.);
    double lir = lid > 1e-10 ? li / lid : (li + 1.) / (lid + 1.);
    int n = (int)nodes.size();
    std::printf("Interior & Boundary l1 error (%d nodes):   Absolute = %le,    Relative = %le \n", n, l1, l1r);
    std::printf("Interior & Boundary l2 error (%d nodes):   Absolute = %le,    Relative = %le \n", n, l2, l2r);
    std::printf("Interior & Boundary linf error (%d nodes): Absolute = %le,    Relative = %le \n", n, li, lir);
    if (l2r > rel_err_max) {
        std::printf("[PDE] Error! l2 relative error (=%f%%) is too high to continue. We require %f%% or less.\n", 100. * l2r,
                    100. * rel_err_max);
        if (!ignore) std::exit(EXIT_FAILURE);
    }
    if (l1 != l1 || l2 != l2 || li != li) {
        std::cout << "Caught NaNs in error!\n";
        std::exit(EXIT_FAILURE);
    }
    return {l1, l1r, l2, l2r, li, lir};
}

CosineBell_HIP::CosineBell_HIP(Domain* grid, RBFFD_HIP* der, Communicator* comm, double earth_radius, double velocity_angle,
                               double one_revolution_in_seconds, int useHV, bool weightsComputed_)
    : TimeDependentPDE_HIP(grid, der, comm, 0, weightsComputed_), useHyperviscosity(useHV), a(earth_radius),
      alpha(velocity_angle), time_for_revolution(one_revolution_in_seconds) {
    R = a / 3.;
    u0 = (2. * M_PI * a) / time_for_revolution;
}

CosineBell_HIP::~CosineBell_HIP() {
    rbfx_dev_free(der_ref.getContext(), d_cl);
    rbfx_dev_free(der_ref.getContext(), d_ct);
}

RBFFD::DerTypes CosineBell_HIP::rhsOperators() {
    return RBFFD::LAMBDA | RBFFD::THETA | (useHyperviscosity ? RBFFD::HV : 0u);
}

void CosineBell_HIP::rhsCoefficients(std::vector<const double*>& c) {
    rbfx_ctx* ctx = der_ref.getContext();
    int n = (int)grid_ref.getNodeListSize();
    if (!d_cl) {
        RBFX_CALL(ctx, rbfx_dev_alloc(ctx, (size_t)n * 8, (void**)&d_cl));
        RBFX_CALL(ctx, rbfx_dev_alloc(ctx, (size_t)n * 8, (void**)&d_ct));
        RBFX_CALL(ctx, rbfx_cosine_bell_coefs(ctx, n, der_ref.getGPUNodes(), a, alpha, time_for_revolution, d_cl, d_ct));
    }
    // ascending bit order: HV (0x20) < LAMBDA (0x40) < THETA (0x80)
    c.clear();
    if (useHyperviscosity) c.push_back(nullptr);
    c.push_back(d_cl);
    c.push_back(d_ct);
}

}  // namespace rbfx
