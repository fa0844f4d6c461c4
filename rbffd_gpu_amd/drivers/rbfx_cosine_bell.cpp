// Cosine-bell advection driver (tests/experiments/cosine_bell_vcl/main.cpp and test.conf, on the
// GPU classes): solid-body rotation of the bell of radius 1/3 centred at (3pi/2, 0) with flow
// angle alpha, 12 days per revolution, RK4, optional Delta^k hyperviscosity scaled
// -gamma N^-k. Nodes: icosahedral level L (N = 10*4^L + 2) or the seeded spiral.
// Prints the reference's norm lines, then one JSON line with the errors and timings.
// The nodes are renumbered along a Morton curve before the stencils are built (numbering
// "morton", the default) or kept in generator order ("input").
//
// usage: rbfx_cosine_bell <ico:L|fib:N> <n> <eps_c1> <eps_c2> <hv_k> <hv_gamma> <steps_per_rev>
//                         <revolutions> <f64|f32> [alpha_denom=2] [dim=2] [numbering=morton]
//                         [stencils=kdtree] [max_steps=0: all] [dump_weights=1 if N <= 100000]
// max_steps stops the run after that many RK4 steps (for step-by-step comparisons);
// dump_weights writes cosine_bell_weights.bin (HV, LAMBDA, THETA in stencil order).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "grid.h"
#include "rbffd_hip.h"
#include "time_dependent_pde_hip.h"

using namespace rbfx;

int main(int argc, char** argv) {
    if (argc < 10) {
        std::fprintf(stderr,
                     "usage: %s ico:L|fib:N n c1 c2 hv_k hv_gamma steps_per_rev revs f64|f32 [alpha_denom] [dim] "
                     "[morton|input] [kdtree|gpu] [max_steps] [dump_weights]\n",
                     argv[0]);
        return 2;
    }
    std::string src = argv[1];
    std::vector<double> xyz;
    if (src.rfind("ico:", 0) == 0) {
        int32_t N = 0;
        int lev = std::atoi(src.c_str() + 4);
        rbfx_nodes_icosahedral(lev, nullptr, &N);
        xyz.resize(3 * (size_t)N);
        rbfx_nodes_icosahedral(lev, xyz.data(), &N);
    } else {
        int N = std::atoi(src.c_str() + 4);
        xyz.resize(3 * (size_t)N);
        rbfx_nodes_fibonacci(N, 0.0, 0, xyz.data());
    }
    size_t N = xyz.size() / 3;
    std::vector<NodeType> nodes(N);
    for (size_t i = 0; i < N; i++) nodes[i] = NodeType(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    int n = std::atoi(argv[2]);
    double c1 = std::atof(argv[3]), c2 = std::atof(argv[4]);
    int hv_k = std::atoi(argv[5]);
    double hv_gamma = std::atof(argv[6]);
    int steps = std::atoi(argv[7]), revs = std::atoi(argv[8]);
    bool f32 = std::string(argv[9]) == "f32";
    double alpha_denom = argc > 10 ? std::atof(argv[10]) : 2.0;
    int dim = argc > 11 ? std::atoi(argv[11]) : 2;
    const bool morton = argc > 12 ? std::string(argv[12]) != "input" : true;
    const Grid::st_generator_t stencils =
        argc > 13 && std::string(argv[13]) == "gpu" ? Grid::ST_KDTREE_GPU : Grid::ST_KDTREE;
    const long max_steps = argc > 14 ? std::atol(argv[14]) : 0;
    const double T = 1036800.0, a = 1.0;
    double alpha = alpha_denom != 0 ? M_PI / alpha_denom : 0.0;

    auto t0 = std::chrono::steady_clock::now();
    Domain grid{Grid(nodes)};
    if (morton) grid.sortNodesMorton();
    grid.generateStencils((unsigned)n, stencils);
    auto t1 = std::chrono::steady_clock::now();
    RBFFD_HIP der(RBFFD::LAMBDA | RBFFD::THETA | RBFFD::HV, &grid, dim, 0);
    int useHV = hv_gamma != 0.0;
    der.setUseHyperviscosity(useHV);
    der.setEpsilonByParameters(c1, c2);
    der.setHVScalars(hv_k, hv_gamma);
    der.computeAllWeightsForAllStencils();
    auto t2 = std::chrono::steady_clock::now();
    Communicator comm;
    CosineBell_HIP pde(&grid, &der, &comm, a, alpha, T, useHV, true);
    if (f32) pde.setPrecision(RBFX_F32, RBFX_F64);
    ExactAdvection exact(a);
    pde.setStartEndTime(0., T * revs);
    pde.fillInitialConditions(&exact);
    double dt = T / steps;
    rbfx_sync(der.getContext());
    auto t3 = std::chrono::steady_clock::now();
    long done = 0;
    for (int r = 0; r < revs; r++) {
        const long todo = max_steps > 0 ? std::min<long>(steps, max_steps - done) : steps;
        if (todo <= 0) break;
        pde.advanceSteps(dt, (int)todo);
        done += todo;
    }
    rbfx_sync(der.getContext());
    auto t4 = std::chrono::steady_clock::now();
    std::printf("============== Verifying Accuracy of Final Solution =============\n");
    std::vector<double> nr = pde.checkGlobalError(&exact, 1.0, true);
    double be = 0;
    const std::vector<double>& dg = der.getSolveDiagnostics();
    for (size_t s = 0; s < N; s++) be = std::fmax(be, dg[4 * s]);
    auto sec = [](auto a_, auto b_) { return std::chrono::duration<double>(b_ - a_).count(); };
    std::vector<double>& U = pde.getSolution();
    FILE* o = std::fopen("cosine_bell_solution.bin", "wb");
    if (o) {
        std::fwrite(U.data(), 8, U.size(), o);
        std::fclose(o);
    }
    const bool dump = argc > 15 ? std::atoi(argv[15]) != 0 : N <= 100000;
    if (dump) {  // weights of HV, LAMBDA, THETA in stencil order, for checking runs
        FILE* w = std::fopen("cosine_bell_weights.bin", "wb");
        for (RBFFD::DerType t : {RBFFD::HV, RBFFD::LAMBDA, RBFFD::THETA})
            for (size_t s = 0; s < N; s++)
                if (der.getStencilWeights(t, (int)s)) std::fwrite(der.getStencilWeights(t, (int)s), 8, (size_t)n, w);
        std::fclose(w);
    }
    std::printf("{\"N\": %zu, \"n\": %d, \"eps\": %.6g, \"hv_k\": %d, \"hv_gamma\": %g, \"dim\": %d, \"steps\": %d, "
                "\"revs\": %d, \"precision\": \"%s\", \"dt\": %.17g, \"l1rel\": %.17g, \"l2\": %.17g, \"l2rel\": %.17g, "
                "\"linf\": %.17g, \"linfrel\": %.17g, \"max_backward_error\": %.3e, \"stencil_s\": %.3f, "
                "\"weights_s\": %.3f, \"step_s\": %.4f, \"steps_per_s\": %.1f}\n",
                N, n, der.getEpsilon()[0], hv_k, hv_gamma, dim, steps, revs, f32 ? "f32" : "f64", dt, nr[1], nr[2], nr[3],
                nr[4], nr[5], be, sec(t0, t1), sec(t1, t2), sec(t3, t4), done / sec(t3, t4));
    return 0;
}
