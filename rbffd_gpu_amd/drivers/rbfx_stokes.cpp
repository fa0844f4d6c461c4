// Steady Stokes driver: the C++ drop-in classes end to end (Grid -> RBFFD_HIP ->
// StokesSteadyPDE_HIP), like the reference's tests/util/viennacl_gmres_poisson/main_stokes.cpp:
// icosahedral nodes (Morton order), n-point kd-tree stencils, eps = c1 sqrt(N) - c2, the
// LSFC / XSFC / YSFC / ZSFC weights on the GPU, then GMRES(krylov) with the velocity-block
// ILU(0) on a body force of spherical harmonics. Prints one JSON line; writes the nodes
// (N x 3), stencils (N x n int32), right-hand side and solution (4N + 4 each) to
// <out_prefix>_{nodes,stencils,rhs,solution}.bin.
//
// usage: rbfx_stokes ico:<level> <n> <c1> <c2> <tol> <max_iter> <krylov> <f64|f32> <out_prefix>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "grid.h"
#include "rbffd_hip.h"
#include "stokes_steady_pde_hip.h"

using namespace rbfx;

int main(int argc, char** argv) {
    if (argc != 10 || std::string(argv[1]).rfind("ico:", 0) != 0) {
        std::fprintf(stderr, "usage: %s ico:<level> n c1 c2 tol max_iter krylov f64|f32 out_prefix\n", argv[0]);
        return 2;
    }
    const int lev = std::atoi(argv[1] + 4);
    int32_t N0 = 0;
    if (rbfx_nodes_icosahedral(lev, nullptr, &N0) != RBFX_OK) return 3;
    std::vector<double> xyz(3 * (size_t)N0);
    rbfx_nodes_icosahedral(lev, xyz.data(), &N0);
    std::vector<NodeType> nodes(N0);
    for (int i = 0; i < N0; i++) nodes[i] = NodeType(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    const int n = std::atoi(argv[2]);
    const double c1 = std::atof(argv[3]), c2 = std::atof(argv[4]), tol = std::atof(argv[5]);
    const int max_iter = std::atoi(argv[6]), krylov = std::atoi(argv[7]);
    const int prec = std::string(argv[8]) == "f32" ? RBFX_F32 : RBFX_F64;
    const std::string out = argv[9];

    Domain grid{Grid(nodes)};
    grid.sortNodesMorton();
    grid.generateStencils((unsigned)n, Grid::ST_KDTREE);
    RBFFD_HIP der(RBFFD::LSFC | RBFFD::XSFC | RBFFD::YSFC | RBFFD::ZSFC, &grid, 3, 0);
    der.setEpsilonByParameters(c1, c2);
    der.computeAllWeightsForAllStencils();
    Communicator comm;
    StokesSteadyPDE_HIP pde(&grid, &der, &comm);
    pde.assemble();

    // body force: a few real spherical harmonics per velocity component, zero divergence target
    const size_t N = grid.getNodeListSize();
    std::vector<double> pts(3 * N);
    for (size_t i = 0; i < N; i++) {
        const NodeType& p = grid.getNode((int)i);
        pts[3 * i] = p.x();
        pts[3 * i + 1] = p.y();
        pts[3 * i + 2] = p.z();
    }
    const int lm[3][2][2] = {{{2, 1}, {5, -3}}, {{3, 2}, {9, 4}}, {{1, 0}, {6, -5}}};
    const double coef[2] = {1.0, 0.5};
    std::vector<double> F(4 * N + 4, 0.0), y(N);
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 2; k++) {
            if (rbfx_sph_harm(lm[c][k][0], lm[c][k][1], (int32_t)N, pts.data(), y.data()) != RBFX_OK) return 3;
            for (size_t i = 0; i < N; i++) F[c * N + i] += coef[k] * y[i];
        }
    pde.setRHS(F);
    pde.setSolverParams(tol, max_iter, krylov, prec);
    pde.solve();
    std::vector<double>& U = pde.getSolution();
    std::vector<double> AU = pde.applyOperator(U);
    double rn = 0, bn = 0;
    for (size_t i = 0; i < F.size(); i++) {
        rn += (F[i] - AU[i]) * (F[i] - AU[i]);
        bn += F[i] * F[i];
    }
    std::vector<int32_t> st(N * (size_t)n);
    for (size_t i = 0; i < N; i++)
        for (int j = 0; j < n; j++) st[i * n + j] = grid.getStencil((int)i)[j];
    auto write = [&](const std::string& name, const void* p, size_t bytes) {
        FILE* f = std::fopen((out + "_" + name + ".bin").c_str(), "wb");
        if (!f || std::fwrite(p, 1, bytes, f) != bytes) std::exit(4);
        std::fclose(f);
    };
    write("nodes", pts.data(), pts.size() * 8);
    write("stencils", st.data(), st.size() * 4);
    write("rhs", F.data(), F.size() * 8);
    write("solution", U.data(), U.size() * 8);
    std::printf("{\"N\": %zu, \"n\": %d, \"eps\": %.17g, \"iterations\": %d, \"error\": %.6e, \"true_rel_residual\": %.6e}\n",
                N, n, c1 * std::sqrt((double)N) - c2, pde.iterations(), pde.error(), std::sqrt(rn / bn));
    return 0;
}
