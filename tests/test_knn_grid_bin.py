"""Host side of the GPU kd-tree stencil search: the uniform-grid binning (rbfx_knn_grid_bin in
csrc/rbfx_knn.cpp, OpenMP) that sorts the nodes by cell before the search kernel runs. A small
C++ driver is compiled against the built object and checks, on the CPU, that every node lands
exactly once in the cell its coordinates select, that the cell table is a monotone prefix sum
ending at N, and that the sorted coordinates are the nodes' own."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ = os.path.join(ROOT, "rbffd_gpu_amd", "_build", "rbfx_knn.o")

DRIVER = r'''
#include <cstdio>
#include <cstdint>
#include <vector>
#include "rbfx_host_internal.h"
extern "C" int rbfx_nodes_fibonacci(int32_t n, double jitter, uint64_t seed, double* xyz);
int main() {
    int bad = 0;
    for (int N : {1, 7, 1000, 100003, 400000}) {
        std::vector<double> x(3 * (size_t)N);
        rbfx_nodes_fibonacci(N, 1e-3, 0, x.data());
        if (N == 7) for (int i = 0; i < 7; i++) { x[3 * i] = x[0]; x[3 * i + 1] = x[1]; x[3 * i + 2] = x[2]; }
        double lo[3], h; int32_t g[3];
        std::vector<int32_t> cs, pid; std::vector<double> pts;
        if (rbfx_knn_grid_bin(N, x.data(), 50, lo, &h, g, cs, pts, pid) != 0) { bad++; continue; }
        const size_t nc = (size_t)g[0] * g[1] * g[2];
        bool ok = cs.size() == nc + 1 && cs[0] == 0 && cs[nc] == N && pid.size() == (size_t)N;
        std::vector<char> seen(N, 0);
        for (size_t c = 0; ok && c < nc; c++) {
            ok = cs[c + 1] >= cs[c];
            for (int s = cs[c]; ok && s < cs[c + 1]; s++) {
                const int i = pid[s];
                ok = i >= 0 && i < N && !seen[i];
                if (!ok) break;
                seen[i] = 1;
                size_t cc = 0;
                for (int d = 2; d >= 0; d--) {
                    int cd = (int)((x[3 * (size_t)i + d] - lo[d]) * (1.0 / h));
                    cd = cd < 0 ? 0 : (cd >= g[d] ? g[d] - 1 : cd);
                    cc = cc * g[d] + cd;
                }
                ok = cc == c && pts[3 * (size_t)s] == x[3 * (size_t)i] && pts[3 * (size_t)s + 1] == x[3 * (size_t)i + 1] &&
                     pts[3 * (size_t)s + 2] == x[3 * (size_t)i + 2];
            }
        }
        for (int i = 0; ok && i < N; i++) ok = seen[i];
        std::printf("N %d grid %d %d %d: %s\n", N, g[0], g[1], g[2], ok ? "ok" : "FAIL");
        bad += !ok;
    }
    std::vector<double> nan_x(6, 0.0);
    nan_x[4] = 0.0 / 0.0;
    double lo[3], h; int32_t g[3];
    std::vector<int32_t> cs, pid; std::vector<double> pts;
    if (rbfx_knn_grid_bin(2, nan_x.data(), 1, lo, &h, g, cs, pts, pid) == 0) { std::printf("NaN accepted\n"); bad++; }
    return bad;
}
'''


def test_grid_binning_host(tmp_path):
    if not os.path.exists(OBJ):
        pytest.skip("build first: make -C rbffd_gpu_amd all")
    src = tmp_path / "bin_driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "bin_driver"
    subprocess.run(["g++", "-O2", "-fopenmp", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "rbffd_gpu_amd", "csrc"), str(src), OBJ, "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
