// Stencil generator driver, the counterpart of the reference's tests/unified/stencil_generator.cxx:
// reads a node file with GridReader, loads the stencil file of that grid if one is present in
// the working directory (Grid::loadStencilsFromFile), and otherwise generates the kd-tree
// stencils and writes them in the reference's layout (Grid::writeStencilsToFile,
// "stencils_maxsz<n>_<grid file name>", lines "<size> i0 i1 ...").
// Prints one JSON line, and writes the stencils' average / min / max radii
// (Grid::computeStencilRadii) as 3*S float64 to radii_<grid file name>.bin.
//
// usage: rbfx_stencil_generator <grid_file> <num_cols> <grid_size> <stencil_size> [kdtree|gpu]
//   num_cols: columns per line of the node file (x y z first, the rest skipped)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "grid.h"

using namespace rbfx;

int main(int argc, char** argv) {
    if (argc != 5 && argc != 6) {
        std::fprintf(stderr, "usage: %s grid_file num_cols grid_size stencil_size [kdtree|gpu]\n", argv[0]);
        return 2;
    }
    std::string path = argv[1];
    int num_cols = std::atoi(argv[2]);
    long grid_size = std::atol(argv[3]);
    int n = std::atoi(argv[4]);
    std::string gen = argc == 6 ? argv[5] : "kdtree";
    if (num_cols < 3 || n < 1 || (gen != "kdtree" && gen != "gpu")) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    std::string name = path.substr(path.find_last_of('/') + 1);

    GridReader grid(path, num_cols - 3, grid_size > 0 ? (unsigned)grid_size : (unsigned)-1);
    grid.setMaxStencilSize((unsigned)n);
    grid.generate();
    bool generated = false;
    if (grid.loadStencilsFromFile(name) != Grid::GRID_AND_STENCILS_LOADED) {
        grid.generateStencils(gen == "gpu" ? Grid::ST_KDTREE_GPU : Grid::ST_KDTREE);
        grid.writeStencilsToFile(name);
        generated = true;
    }
    const size_t S = grid.getStencilsSize();
    std::vector<double> radii(3 * S);
    for (size_t i = 0; i < S; i++) {
        radii[i] = grid.getStencilRadius((int)i);
        radii[S + i] = grid.getMinStencilRadius((int)i);
        radii[2 * S + i] = grid.getMaxStencilRadius((int)i);
    }
    std::string rn = "radii_" + name + ".bin";
    FILE* f = std::fopen(rn.c_str(), "wb");
    if (!f || std::fwrite(radii.data(), 8, radii.size(), f) != radii.size()) return 3;
    std::fclose(f);
    std::printf("{\"nodes\": %u, \"stencils\": %zu, \"max_size\": %u, \"generated\": %s, \"file\": \"%s\"}\n",
                grid.getNodeListSize(), S, grid.getMaxStencilSize(), generated ? "true" : "false",
                grid.stencilFilename(name).c_str());
    return 0;
}
