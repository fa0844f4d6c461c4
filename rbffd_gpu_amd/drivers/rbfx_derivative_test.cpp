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
    // weight files: X as MatrixMarket text (.mtx), Y as the binary form (.bmtx)
    der.setAsciiWeights(1);
    der.writeToFile(RBFFD::X, out + "_x.mtx");
    der.setAsciiWeights(0);
    der.writeToFile(RBFFD::Y, out + "_y.bmtx");
    // read back into a second operator object on the same grid; its derivatives must equal the
    // first object's bit for bit (%24.16le keeps 17 significant digits: exact for doubles)
    RBFFD_HIP der2(RBFFD::X | RBFFD::Y, &grid, 3, 0);
    der2.setAsciiWeights(1);
    int rc_x = der2.loadFromFile(RBFFD::X, out + "_x.mtx");
    der2.setAsciiWeights(0);
    int rc_y = der2.loadFromFile(RBFFD::Y, out + "_y.bmtx");
    int rc_missing = der2.loadFromFile(RBFFD::Y, out + "_missing.bmtx");
    std::vector<double> dx2, dy2;
    der2.applyWeightsForDeriv(RBFFD::X, u, dx2, true);
    der2.applyWeightsForDeriv(RBFFD::Y, u, dy2, false);
    bool reload_identical = dx2 == dx && dy2 == dy;
    // the pointer overload writes deriv[i] for i in [start, start + nb) only
    const unsigned start = (unsigned)(N / 3), nb = (unsigned)(N / 5);
    std::vector<double> part(N, -7.0);
    der.applyWeightsForDeriv(RBFFD::X, start, nb, u.data(), part.data(), true);
    bool partial_ok = true;
    for (size_t i = 0; i < N; i++) partial_ok &= (i >= start && i < start + nb) ? part[i] == dx[i] : part[i] == -7.0;
    std::printf("{\"N\": %zu, \"n\": %d, \"eps\": %g, \"poly_deg\": %d, \"max_backward_error\": %.3e, "
                "\"rel_err_vs_analytic\": {\"x\": %.3e, \"y\": %.3e, \"z\": %.3e}, "
                "\"reload_identical\": %s, \"load_codes\": [%d, %d, %d], \"partial_apply_ok\": %s}\n",
                N, n, eps, deg, be, relmax(dx, fx), relmax(dy, fy), relmax(dz, fz), reload_identical ? "true" : "false",
                rc_x, rc_y, rc_missing, partial_ok ? "true" : "false");
    return 0;
}
