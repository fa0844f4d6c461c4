// Time-dependent PDE driver on the GPU: the analogue of TimeDependentPDE_CL / _VCL
// (src/pdes/time_dependent_pde_cl.h:78-190, time_dependent_pde_vcl.h:63-190) and of the
// CosineBell family (src/pdes/hyperbolic/cosine_bell{,_cl,_vcl}.h).
//
// The reference plugs the PDE right-hand side in as an OpenCL device function (cosine_bell_solve.cl)
// or a virtual ViennaCL solve(); here a PDE subclass names its operators and one coefficient
// array per operator, f = sum_o c_o (D_o u), and rbfx_rk4 fuses operators, right-hand side and
// stage update into one kernel per RK4 stage, replayed from a hipGraph.
#pragma once
#include <cmath>
#include <string>
#include <vector>

#include "grid.h"
#include "rbffd_hip.h"

namespace rbfx {

// ExactSolution (src/exact_solutions/exact_solution.h:22-27)
class ExactSolution {
  public:
    explicit ExactSolution(int dimension) : dim_num(dimension) {}
    virtual ~ExactSolution() {}
    virtual double operator()(double x, double y, double z, double t = 0.) = 0;
    virtual double at(const NodeType& r, double t = 0.) { return (*this)(r.x(), r.y(), r.z(), t); }

  protected:
    int dim_num;
};

// ExactAdvection (tests/experiments/cosine_bell_vcl/exact_advection.h:29-54): the cosine bell of
// radius a/3 centred at (3pi/2, 0); time-independent (one revolution returns it to the start).
class ExactAdvection : public ExactSolution {
  public:
    explicit ExactAdvection(double sphere_radius) : ExactSolution(3), a(sphere_radius) {}
    double operator()(double x, double y, double z, double t = 0.) override;

  private:
    double a;
};

class TimeDependentPDE_HIP {
  public:
    enum TimeScheme { EULER = 0, MIDPOINT, RK4 };

    TimeDependentPDE_HIP(Domain* grid, RBFFD_HIP* der, Communicator* comm, int gpu_type = 0, bool weightsComputed = false);
    virtual ~TimeDependentPDE_HIP();

    virtual void fillInitialConditions(ExactSolution* exact = nullptr);
    virtual void advance(TimeScheme which, double delta_t);
    void advanceSteps(double delta_t, int n_steps);  // n steps in one call (graph replays)
    virtual void assemble();                          // computes weights if not done yet
    void syncCPUtoGPU();                              // device state -> U_G
    std::vector<double>& getSolution() { return U_G; }
    double getTime() const { return cur_time; }
    void setTime(double t) { cur_time = t; }
    void setStartEndTime(double s, double e) {
        start_time = cur_time = s;
        end_time = e;
    }
    // Precision of the stepped fields: RBFX_F64, or RBFX_F32 storage of weights and stage
    // fields with FP64 accumulation and FP64 state.
    void setPrecision(int field_storage_, int accum_) {
        field_storage = field_storage_;
        accum = accum_;
    }
    // PDE::calcSolNorms (src/pdes/pde.cpp:463-533) against the exact solution at the current time:
    // out = {l1, l1rel, l2, l2rel, linf, linfrel}; prints like the reference; exits on NaN.
    std::vector<double> checkGlobalError(ExactSolution* exact, double rel_err_max = 1.0, bool ignore_high_error = true);
    rbfx_rk4* getStepper() { return rk; }

  protected:
    // The right-hand side f = sum_o c_o (D_o u): operator DerTypes (ascending bits) and one
    // device coefficient array per operator (nullptr = 1).
    virtual RBFFD::DerTypes rhsOperators() = 0;
    virtual void rhsCoefficients(std::vector<const double*>& d_coef) = 0;
    void ensureStepper();

    Domain& grid_ref;
    RBFFD_HIP& der_ref;
    int gpu_type;
    bool weightsComputed;
    double cur_time = 0., start_time = 0., end_time = 0.;
    std::vector<double> U_G;
    ExactSolution* exact_ptr = nullptr;
    int field_storage = RBFX_F64, accum = RBFX_F64;
    rbfx_rk4* rk = nullptr;
    double* d_state = nullptr;
    bool state_on_device = false;
};

// CosineBell (src/pdes/hyperbolic/cosine_bell.{h,cpp}) on the GPU: operators LAMBDA, THETA and,
// with hyperviscosity, HV; coefficients -(u_vel/a), -(v_vel/a), 1.
class CosineBell_HIP : public TimeDependentPDE_HIP {
  public:
    CosineBell_HIP(Domain* grid, RBFFD_HIP* der, Communicator* comm, double earth_radius, double velocity_angle,
                   double one_revolution_in_seconds, int useHyperviscosity, bool weightsComputed = false);
    ~CosineBell_HIP() override;
    double getMaxVelocity(double) const { return u0; }

  protected:
    RBFFD::DerTypes rhsOperators() override;
    void rhsCoefficients(std::vector<const double*>& d_coef) override;

  private:
    int useHyperviscosity;
    double a, alpha, time_for_revolution, R, u0;
    double* d_cl = nullptr;
    double* d_ct = nullptr;
};

}  // namespace rbfx
