// Derivative test driver (config: N nodes on the sphere, n-point kd-tree stencils, Gaussian RBF
// with fixed eps, polynomial degree d): the C++ drop-in classes end to end, like the
// reference's test executables (tests/unified/compute_weights.cxx + evaluate_derivatives.cxx).
// Weights of X, Y, Z and the Laplace-Beltrami operator are computed on the GPU and applied to
// f = exp(x) sin(2y) + z^3; errors against the analytic derivatives are printed as JSON and the
// derivatives are written to <out_prefix>.bin for comparison with the CPU restatement.
//
// usage: rbfx_derivative_test <nodes.bin|fib:N> <n> <eps> <poly_deg> <out_prefix>
//   nodes.bin: int64 N then N*3 float64; fib:N = seeded spiral +-1e-3 (seed 0)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "grid.h"
#include "rbffd_hip.h"

using namespace rbfx;

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s nodes.bin|fib:N n eps poly_deg out_prefix\n", argv[0]);
        return 2;
    }
    std::vector<NodeType> nodes;
    std::string src = argv[1];
    if (src.rfind("fib:", 0) == 0) {
        int N = std::atoi(src.c_str() + 4);
        std::vector<double> xyz(3 * (size_t)N);
        if (rbfx_nodes_fibonacci(N, 1e-3, 0, xyz.data()) != RBFX_OK) return 3;
        for (int i = 0; i < N; i++) nodes.push_back(NodeType(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    } else {
        FILE* f = std::fopen(src.c_str(), "rb");
        if (!f) return 3;
        int64_t N = 0;
        if (std::fread(&N, 8, 1, f) != 1) return 3;
        std::vector<double> xyz(3 * (size_t)N);
        if (std::fread(xyz.data(), 8, xyz.size(), f) != xyz.size()) return 3;
        std::fclose(f);
        for (int64_t i = 0; i < N; i++) nodes.push_back(NodeType(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    }
    int n = std::atoi(argv[2]);
    double eps = std::atof(argv[3]);
    int deg = std::atoi(argv[4]);
    std::string out = argv[5];

    Grid grid(nodes);
    grid.generateStencils((unsigned)n, Grid::ST_KDTREE);
    RBFFD_HIP der(RBFFD::X | RBFFD::Y | RBFFD::Z | RBFFD::LSFC, &grid, 3, 0);
    der.setEpsilon(eps);
    der.setPolynomialDegree(deg);
    der.computeAllWeightsForAllStencils();

    size_t N = nodes.size();
    std::vector<double> u(N), fx(N), fy(N), fz(N);
    for (size_t i = 0; i < N; i++) {
        double x = nodes[i].x(), y = nodes[i].y(), z = nodes[i].z();
        u[i] = std::exp(x) * std::sin(2 * y) + z * z * z;
        fx[i] = std::exp(x) * std::sin(2 * y);
        fy[i] = 2 * std::exp(x) * std::cos(2 * y);
        fz[i] = 3 * z * z;
    }
    std::vector<double> dx, dy, dz, dl;
    der.applyWeightsForDeriv(RBFFD::X, u, dx, true);
    der.applyWeightsForDeriv(RBFFD::Y, u, dy, false);
    der.applyWeightsForDeriv(RBFFD::Z, u, dz, false);
    der.applyWeightsForDeriv(RBFFD::LSFC, u, dl, false);
    auto relmax = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double e = 0, m = 0;
        for (size_t i = 0; i < N; i++) {
            e = std::fmax(e, std::fabs(a[i] - b[i]));
            m = std::fmax(m, std::fabs(b[i]));
        }
        return e / m;
    };
    double be = 0;
    const std::vector<double>& dg = der.getSolveDiagnostics();
    for (size_t s = 0; s < N; s++) be = std::fmax(be, dg[4 * s]);
    FILE* o = std::fopen((out + ".bin").c_str(), "wb");
    for (auto* v : {&dx, &dy, &dz, &dl}) std::fwrite(v->data(), 8, N, o);
    // weights of the four operators, stencil order (N x n each)
    for (RBFFD::DerType t : {RBFFD::X, RBFFD::Y, RBFFD::Z, RBFFD::LSFC})
        for (size_t s = 0; s < N; s++) std::fwrite(der.getStencilWeights(t, (int)s), 8, (size_t)n, o);
    std::fclose(o);
    der.setAsciiWeights(1);
    der.writeToFile(RBFFD::X, out + "_x.mtx");
    std::printf("{\"N\": %zu, \"n\": %d, \"eps\": %g, \"poly_deg\": %d, \"max_backward_error\": %.3e, "
                "\"rel_err_vs_analytic\": {\"x\": %.3e, \"y\": %.3e, \"z\": %.3e}}\n",
                N, n, eps, deg, be, relmax(dx, fx), relmax(dy, fy), relmax(dz, fz));
    return 0;
}
