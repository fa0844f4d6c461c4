// Test infrastructure (oracle side): a small driver that links the reference's own
// Grid::generateStencilsKDTree (src/grids/grid_interface.cpp:688-725) and KDTree
// (external/kdtree/KDTree.h) compiled from the sources under the reference tree, so that
// kd-tree stencil lists can be produced by the reference implementation itself.
//
// Usage: ref_knn <nodes.bin> <k> <out.bin> [grid_name]
//   nodes.bin : int64 N, then N*3 float64 (x,y,z per node)
//   out.bin   : int32 N, int32 k, then N*k int32 stencil indices (row 0 = centre)
//   grid_name : also write the reference's stencil file (Grid::writeStencilsToFile,
//               grid_interface.cpp:144-170) for that grid name into the working directory
//
// Only tests/, bench.py's CPU baseline and the smoke check use this binary, as a checker.
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <vector>

#include "grids/grid_interface.h"

int main(int argc, char** argv) {
    if (argc != 4 && argc != 5) {
        std::fprintf(stderr, "usage: %s nodes.bin k out.bin [grid_name]\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) { std::perror("open nodes"); return 2; }
    int64_t n = 0;
    if (std::fread(&n, sizeof(n), 1, f) != 1 || n <= 0) { std::fprintf(stderr, "bad header\n"); return 2; }
    std::vector<double> xyz(3 * (size_t)n);
    if (std::fread(xyz.data(), sizeof(double), xyz.size(), f) != xyz.size()) {
        std::fprintf(stderr, "short nodes file\n");
        return 2;
    }
    std::fclose(f);
    unsigned int k = (unsigned int)std::atoi(argv[2]);

    std::vector<NodeType> nodes;
    nodes.reserve(n);
    for (int64_t i = 0; i < n; i++) nodes.push_back(NodeType(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));

    Grid grid(nodes);
    grid.generateStencils(k, Grid::ST_KDTREE);
    if (argc == 5) grid.writeStencilsToFile(argv[4]);

    std::vector<StencilType>& st = grid.getStencils();
    FILE* o = std::fopen(argv[3], "wb");
    if (!o) { std::perror("open out"); return 2; }
    int32_t hdr[2] = {(int32_t)st.size(), (int32_t)grid.getMaxStencilSize()};
    std::fwrite(hdr, sizeof(int32_t), 2, o);
    for (size_t i = 0; i < st.size(); i++) {
        if (st[i].size() != (size_t)hdr[1]) { std::fprintf(stderr, "ragged stencil %zu\n", i); return 3; }
        std::vector<int32_t> row(st[i].begin(), st[i].end());
        std::fwrite(row.data(), sizeof(int32_t), row.size(), o);
    }
    std::fclose(o);
    return 0;
}
