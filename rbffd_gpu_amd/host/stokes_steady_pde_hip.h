// Steady Stokes flow on the sphere on the GPU: the analogue of StokesSteadyPDE
// (src/pdes/elliptic/stokes_steady_pde.{h,cpp}) with the GMRES + ILU(0) solve that the
// reference's working driver runs (tests/util/viennacl_gmres_poisson/main_stokes.cpp:71-110).
//
// The coupled system keeps the reference's block layout (u, v, w, p, then the four constraint
// multipliers) and is never assembled: rbfx_stokes applies it from the LSFC, XSFC, YSFC, ZSFC
// weights of the RBFFD_HIP object, factors the ILU(0) of its velocity block (ilu0_tag(0, 3N))
// and runs restarted GMRES on the device.
#pragma once
#include <string>
#include <vector>

#include "grid.h"
#include "rbffd_hip.h"

namespace rbfx {

class StokesSteadyPDE_HIP {
  public:
    // der must compute at least LSFC | XSFC | YSFC | ZSFC on fixed-width stencils.
    StokesSteadyPDE_HIP(Domain* grid, RBFFD_HIP* der, Communicator* comm, double eta = 1.);
    virtual ~StokesSteadyPDE_HIP();

    // Weights on the GPU if not computed yet, the operator sets, the ILU(0) factor
    // (StokesSteadyPDE::assemble and ilu0_precond's construction).
    virtual void assemble();
    // GMRES from x = 0 on the right-hand side set by setRHS (viennacl::linalg::solve with a
    // gmres_tag and the ILU(0) preconditioner); the solution is getSolution().
    virtual void solve();
    virtual std::string className() { return "stokes_steady"; }

    // Right-hand side in the block layout, 4N + 4 values (the reference's RHS_continuous).
    void setRHS(const std::vector<double>& F);
    // gmres_tag(tol, max_iterations, krylov_dim); precision RBFX_F64, or RBFX_F32 for FP32
    // inner solves inside FP64 refinement. Defaults: main_stokes.cpp's gmres_tag(1e-6, 1000, 80).
    void setSolverParams(double tol, int max_iterations, int krylov_dim, int precision = RBFX_F64);
    std::vector<double>& getSolution() { return U_computed; }
    // y = A x on the host vectors (the reference's prod(LHS, x)).
    std::vector<double> applyOperator(const std::vector<double>& x);
    int iterations() const { return iters; }
    double error() const { return err; }  // the final ||M^-1 (b - A x)|| / ||b||
    unsigned int numUnknowns() const { return 4 * N + 4; }
    rbfx_stokes* getSolver() { return solver; }

  private:
    Domain& grid_ref;
    RBFFD_HIP& der_ref;
    double eta;
    unsigned int N;
    double tol = 1e-6;
    int max_iter = 1000, krylov = 80, precision = RBFX_F64;
    std::vector<double> RHS, U_computed;
    int iters = 0;
    double err = 0.;
    rbfx_stokes* solver = nullptr;
    double* d_b = nullptr;
    double* d_x = nullptr;
};

}  // namespace rbfx
