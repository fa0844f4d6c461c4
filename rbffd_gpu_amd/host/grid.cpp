#include "grid.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

#include "rbfx.h"

namespace rbfx {

void Grid::checkStencilSize() {
    unsigned int nb_rbf = (unsigned)node_list.size();
    if (max_st_size < 1) {
        std::cout << "[Grid] ERROR! Stencil size must be >= 1" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    if (max_st_size > nb_rbf) {
        std::cout << "[Grid] WARNING! Not enough nodes to reach specified stencil_size (size: " << max_st_size
                  << ")! Using new stencil size: " << nb_rbf << std::endl;
        max_st_size = nb_rbf > 1 ? nb_rbf : 1;
    }
}

std::vector<int32_t> Grid::sortNodesMorton() {
    const size_t n = node_list.size();
    std::vector<double> xyz(3 * n);
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) xyz[3 * i + d] = node_list[i][d];
    std::vector<int32_t> perm(n);
    if (n && rbfx_order_morton((int32_t)n, xyz.data(), perm.data()) != RBFX_OK) {
        std::cout << "[Grid] ERROR: rbfx_order_morton failed" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    std::vector<NodeType> sorted(n);
    for (size_t i = 0; i < n; i++) sorted[i] = node_list[perm[i]];
    node_list.swap(sorted);
    stencil_map.clear();
    avg_stencil_radii.clear();
    min_stencil_radii.clear();
    max_stencil_radii.clear();
    return perm;
}

void Grid::generateStencils(unsigned int st_max_size, st_generator_t generator_choice) {
    max_st_size = st_max_size;
    generateStencils(generator_choice);
}

void Grid::generateStencils(st_generator_t generator_choice) {
    if (generator_choice != ST_KDTREE && generator_choice != ST_KDTREE_GPU) {
        std::cout << "[Grid] ERROR: only the kd-tree stencil generator (ST_KDTREE, ST_KDTREE_GPU) is provided"
                  << std::endl;
        std::exit(EXIT_FAILURE);
    }
    checkStencilSize();
    size_t n = node_list.size();
    std::vector<double> xyz(3 * n);
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) xyz[3 * i + d] = node_list[i][d];
    std::vector<int32_t> st(n * (size_t)max_st_size);
    int rc;
    if (generator_choice == ST_KDTREE_GPU) {
        rbfx_ctx* ctx = nullptr;
        rc = rbfx_create(0, &ctx);
        if (rc == RBFX_OK) {
            rc = rbfx_knn_gpu(ctx, (int32_t)n, xyz.data(), (int32_t)max_st_size, st.data(), nullptr);
            if (rc != RBFX_OK) std::cout << "[Grid] " << rbfx_last_error(ctx) << std::endl;
            rbfx_destroy(ctx);
        }
    } else {
        rc = rbfx_knn_kdtree((int32_t)n, xyz.data(), (int32_t)max_st_size, st.data(), 0);
    }
    if (rc != RBFX_OK) {
        std::cout << "[Grid] ERROR: kd-tree stencil generation failed (" << rc << ")" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    stencil_map.assign(n, StencilType());
    for (size_t i = 0; i < n; i++) {
        stencil_map[i].resize(max_st_size);
        for (unsigned j = 0; j < max_st_size; j++) stencil_map[i][j] = st[i * max_st_size + j];
        // radius trimming against max_st_radius (DBL_MAX by default, as in the reference)
        const NodeType& c = node_list[i];
        for (unsigned j = 0; j < max_st_size; j++) {
            const NodeType& p = node_list[stencil_map[i][j]];
            double dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
            if (!(std::sqrt(dx * dx + dy * dy + dz * dz) < max_st_radius)) {
                stencil_map[i].resize(j);
                break;
            }
        }
    }
    computeStencilRadii();
}

void Grid::computeStencilRadii() {
    size_t nb = stencil_map.size();
    avg_stencil_radii.assign(nb, 0.);
    min_stencil_radii.assign(nb, 0.);
    max_stencil_radii.assign(nb, 0.);
    auto dist = [&](int a, int b) {
        NodeType d(node_list[a][0] - node_list[b][0], node_list[a][1] - node_list[b][1], node_list[a][2] - node_list[b][2]);
        return std::sqrt(d.square());
    };
    for (size_t i = 0; i < nb; i++) {
        StencilType& st = stencil_map[i];
        if (st.size() < 2) continue;
        min_stencil_radii[i] = dist(st[1], st[0]);
        max_stencil_radii[i] = dist(st[st.size() - 1], st[0]);
        double s = 0.;
        for (size_t k = 1; k < st.size(); k++) s += dist(st[k], st[0]);
        avg_stencil_radii[i] = s / (st.size() - 1.);
    }
}

std::string Grid::stencilFilename(const std::string& filename) const {
    std::ostringstream prefix;
    prefix << "stencils_maxsz" << max_st_size << "_" << filename;
    return prefix.str();
}

void Grid::writeStencilsToFile(const std::string& filename) const {
    if (max_st_size == 0) {
        std::cout << "[Grid] \tMax stencil size not set. No stencils to write to disk" << std::endl;
        return;
    }
    const std::string fname = stencilFilename(filename);
    std::ofstream fout(fname);
    if (!fout.is_open()) {
        std::printf("Error opening file to write\n");
        std::exit(EXIT_FAILURE);
    }
    for (const StencilType& st : stencil_map) {
        fout << st.size();
        for (int v : st) fout << " " << v;
        fout << "\n";
    }
    std::cout << "[Grid] \tWrote " << stencil_map.size() << " stencils to \t" << fname << std::endl;
}

Grid::GridLoadErrType Grid::loadStencilsFromFile(const std::string& filename) {
    if (max_st_size == 0) {
        std::cout << "[Grid] \tMax stencil size not set. No stencils to read from disk" << std::endl;
        return GRID_AND_STENCILS_LOADED;
    }
    const std::string fname = stencilFilename(filename);
    std::ifstream fin(fname);
    if (!fin.is_open()) {
        std::printf("Error opening stencil file to read\n");
        return NO_STENCIL_FILES;
    }
    stencil_map.clear();
    size_t loaded = 0;
    // a record is "<size> i0 ... i(size-1)"; a record cut short by the end of the file is
    // dropped, as the reference's loop drops it (it keeps a record only while !eof)
    while (fin.good()) {
        unsigned int st_size = 0;
        if (!(fin >> st_size)) break;
        StencilType st(st_size);
        bool ok = true;
        for (unsigned int i = 0; i < st_size && ok; i++) ok = static_cast<bool>(fin >> st[i]);
        if (!ok) break;
        stencil_map.push_back(st);
        loaded += st.size();
    }
    computeStencilRadii();
    std::cout << "[Grid] \tLoaded " << stencil_map.size() << " stencils, with a total of " << loaded
              << " elements from \t" << fname << std::endl;
    return GRID_AND_STENCILS_LOADED;
}

void GridReader::generate() { readNodeList(extra_); }

int GridReader::readNodeList(int expect_num_extra_dbls_per_line) {
    std::ifstream fin(filename_);
    if (!fin.is_open()) {
        std::printf("Error opening node file to read\n");
        std::exit(EXIT_FAILURE);
    }
    node_list.clear();
    unsigned int i = 0;
    while (fin.good()) {
        double x, y, z, junk;
        fin >> x >> y >> z;
        for (int j = 0; j < expect_num_extra_dbls_per_line; j++) fin >> junk;
        if (!fin.eof() && !fin.fail()) {
            if (i < n_nodes_) {
                node_list.push_back(NodeType(x, y, z));
                i++;
            } else {
                break;
            }
        }
    }
    nb_nodes = (unsigned)node_list.size();
    if (n_nodes_ != (unsigned)-1 && n_nodes_ > 0 && nb_nodes < n_nodes_) {
        std::cout << "[GridReader] \tERROR: Found only " << nb_nodes << " nodes \t" << filename_
                  << ". Check your configuration." << std::endl;
        std::exit(EXIT_FAILURE);
    }
    global_num_nodes = nb_nodes;
    return 0;
}

}  // namespace rbfx
