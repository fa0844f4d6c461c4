#include "stokes_steady_pde_hip.h"

#include <cstdlib>
#include <iostream>

namespace rbfx {

namespace {
[[noreturn]] void die(rbfx_ctx* ctx, const char* what, int rc) {
    std::cout << "[StokesSteadyPDE_HIP] ERROR in " << what << " (" << rc << "): " << rbfx_last_error(ctx) << std::endl;
    std::exit(EXIT_FAILURE);
}
#define RBFX_CALL(ctx, call)                      \
    do {                                          \
        int rc_ = (call);                         \
        if (rc_ != RBFX_OK) die(ctx, #call, rc_); \
    } while (0)
const RBFFD::DerTypes STOKES_TYPES = RBFFD::LSFC | RBFFD::XSFC | RBFFD::YSFC | RBFFD::ZSFC;
}  // namespace

StokesSteadyPDE_HIP::StokesSteadyPDE_HIP(Domain* grid, RBFFD_HIP* der, Communicator*, double eta_)
    : grid_ref(*grid), der_ref(*der), eta(eta_), N(grid->getNodeListSize()) {
    if ((der_ref.getComputedTypes() & STOKES_TYPES) != STOKES_TYPES) {
        std::cout << "[StokesSteadyPDE_HIP] ERROR: the RBFFD object must compute LSFC, XSFC, YSFC and ZSFC"
                  << std::endl;
        std::exit(EXIT_FAILURE);
    }
    RHS.assign(numUnknowns(), 0.0);
    U_computed.assign(numUnknowns(), 0.0);
}

StokesSteadyPDE_HIP::~StokesSteadyPDE_HIP() {
    rbfx_stokes_destroy(solver);
    rbfx_dev_free(der_ref.getContext(), d_b);
    rbfx_dev_free(der_ref.getContext(), d_x);
}

void StokesSteadyPDE_HIP::assemble() {
    std::cout << "Assembling...." << std::endl;
    if (!der_ref.getWeights(RBFFD::LSFC).size() || !der_ref.getWeights(RBFFD::LSFC)[0])
        der_ref.computeAllWeightsForAllStencils();
    rbfx_ctx* ctx = der_ref.getContext();
    rbfx_ell* ops64 = der_ref.getGPUOperatorSet(STOKES_TYPES, RBFX_F64);
    rbfx_ell* ops32 = der_ref.getGPUOperatorSet(STOKES_TYPES, RBFX_F32);
    rbfx_stokes_destroy(solver);
    solver = nullptr;
    RBFX_CALL(ctx, rbfx_stokes_create(ctx, ops64, ops32, eta, &solver));
    if (!d_b) {
        RBFX_CALL(ctx, rbfx_dev_alloc(ctx, numUnknowns() * sizeof(double), (void**)&d_b));
        RBFX_CALL(ctx, rbfx_dev_alloc(ctx, numUnknowns() * sizeof(double), (void**)&d_x));
    }
    std::cout << "Done." << std::endl;
}

void StokesSteadyPDE_HIP::setRHS(const std::vector<double>& F) {
    if (F.size() != numUnknowns()) {
        std::cout << "[StokesSteadyPDE_HIP] ERROR: the right-hand side has " << F.size() << " values, not "
                  << numUnknowns() << std::endl;
        std::exit(EXIT_FAILURE);
    }
    RHS = F;
}

void StokesSteadyPDE_HIP::setSolverParams(double tol_, int max_iter_, int krylov_, int precision_) {
    tol = tol_;
    max_iter = max_iter_;
    krylov = krylov_;
    precision = precision_;
}

void StokesSteadyPDE_HIP::solve() {
    if (!solver) assemble();
    rbfx_ctx* ctx = der_ref.getContext();
    std::cout << "Solving...." << std::endl;
    RBFX_CALL(ctx, rbfx_h2d(ctx, d_b, RHS.data(), numUnknowns() * sizeof(double)));
    int32_t it = 0;
    RBFX_CALL(ctx, rbfx_stokes_gmres(solver, d_b, d_x, tol, max_iter, krylov, precision, &it, &err));
    iters = it;
    RBFX_CALL(ctx, rbfx_d2h(ctx, U_computed.data(), d_x, numUnknowns() * sizeof(double)));
    RBFX_CALL(ctx, rbfx_sync(ctx));
    std::cout << "GMRES Iterations: " << iters << std::endl;
    std::cout << "GMRES Error Estimate: " << err << std::endl;
    std::cout << "GMRES Krylov Dim: " << krylov << std::endl;
    std::cout << "GMRES Max Number of Iterations: " << max_iter << std::endl;
    std::cout << "GMRES Tolerance: " << tol << std::endl;
}

std::vector<double> StokesSteadyPDE_HIP::applyOperator(const std::vector<double>& x) {
    if (!solver) assemble();
    rbfx_ctx* ctx = der_ref.getContext();
    std::vector<double> y(numUnknowns());
    RBFX_CALL(ctx, rbfx_h2d(ctx, d_x, x.data(), numUnknowns() * sizeof(double)));
    RBFX_CALL(ctx, rbfx_stokes_apply(solver, d_x, d_b));
    RBFX_CALL(ctx, rbfx_d2h(ctx, y.data(), d_b, numUnknowns() * sizeof(double)));
    RBFX_CALL(ctx, rbfx_sync(ctx));
    return y;
}

}  // namespace rbfx
