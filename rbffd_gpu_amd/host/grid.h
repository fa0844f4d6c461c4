// Grid / GridReader / Domain / Communicator: the node-set and stencil containers that the
// reference's RBFFD and PDE classes take, with the same method names (src/grids/grid_interface.h,
// src/grids/grid_reader.h, src/grids/domain.h, src/utils/comm/communicator.h). Only the parts
// the hot path touches are provided; Domain and Communicator are single-rank.
#pragma once
#include <cfloat>
#include <cstdint>
#include <string>
#include <vector>

namespace rbfx {

// 3-vector of doubles with the accessors the reference's Vec3d offers (Vec3d.h).
struct NodeType {
    double vec[3];
    NodeType() : vec{0, 0, 0} {}
    NodeType(double x, double y, double z = 0.) : vec{x, y, z} {}
    double x() const { return vec[0]; }
    double y() const { return vec[1]; }
    double z() const { return vec[2]; }
    double& operator[](int i) { return vec[i]; }
    const double& operator[](int i) const { return vec[i]; }
    double square() const { return vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]; }
};
typedef std::vector<int> StencilType;

class Grid {
  public:
    // ST_KDTREE_GPU (not in the reference): the kd-tree's lists from the GPU search
    // rbfx_knn_gpu, bit-identical to ST_KDTREE; it uses device 0 for the call.
    enum st_generator_t { ST_BRUTE_FORCE = 0, ST_KDTREE, ST_HASH, ST_KDTREE_GPU };
    // Return codes of the loaders (grid_interface.h:38).
    enum GridLoadErrType {
        GRID_AND_STENCILS_LOADED = 0, NO_GRID_FILES, NO_BOUNDARY_FILES, NO_NORMAL_FILES, NO_STENCIL_FILES,
        NO_RADII_FILES, NO_EXTRA_FILES, NOT_ENOUGH_NODES_IN_FILE
    };

    Grid() {}
    explicit Grid(unsigned int num_nodes) : nb_nodes(num_nodes), global_num_nodes(num_nodes) {}
    explicit Grid(const std::vector<NodeType>& nodes)
        : nb_nodes((unsigned)nodes.size()), global_num_nodes((unsigned)nodes.size()), node_list(nodes) {}
    virtual ~Grid() {}

    // Grid::generateStencils (grid_interface.cpp:507-536): ST_KDTREE runs the reference's
    // kd-tree algorithm (bit-identical lists) through rbfx_knn_kdtree; ST_KDTREE_GPU gives the
    // same lists through rbfx_knn_gpu. The brute-force and hash generators of the reference
    // are not part of this build and exit with an error.
    void generateStencils(st_generator_t generator_choice = ST_KDTREE);
    void generateStencils(unsigned int st_max_size, st_generator_t generator_choice = ST_KDTREE);
    void computeStencilRadii();  // grid_interface.cpp:539-633 (min/max/avg distance to centre)

    // Renumbers the nodes along a 3-D Morton curve (rbfx_order_morton): node i of the new list
    // is node perm[i] of the old one, and perm is returned. Like the reference's
    // Grid::sortNodes (grid_interface.cpp:432-448, boundary nodes first) this only relabels
    // node_list. With the new numbering the field values of a stencil's nodes share cache lines,
    // which halves the gathered bytes of the derivative kernels. Call it before
    // generateStencils; stencils generated earlier are dropped.
    std::vector<int32_t> sortNodesMorton();

    void setMaxStencilSize(unsigned int s) { max_st_size = s; }
    unsigned int getMaxStencilSize() const { return max_st_size; }
    unsigned int getNodeListSize() const { return (unsigned)node_list.size(); }
    std::vector<NodeType>& getNodeList() { return node_list; }
    NodeType& getNode(unsigned int i) { return node_list[i]; }
    void setNode(unsigned int i, NodeType n) { node_list[i] = n; }
    unsigned int getStencilsSize() const { return (unsigned)stencil_map.size(); }
    unsigned int getStencilSize(int i) const { return (unsigned)stencil_map[i].size(); }
    std::vector<StencilType>& getStencils() { return stencil_map; }
    StencilType& getStencil(int i) { return stencil_map[i]; }
    unsigned int getGlobalNodeListSize() const { return global_num_nodes; }
    std::vector<double>& getStencilRadii() { return avg_stencil_radii; }
    double getStencilRadius(int i) const { return avg_stencil_radii[i]; }
    double getMinStencilRadius(int i) const { return min_stencil_radii[i]; }
    double getMaxStencilRadius(int i) const { return max_stencil_radii[i]; }

    // Stencil files in the reference's layout (Grid::writeStencilsToFile / loadStencilsFromFile,
    // grid_interface.cpp:144-170, 379-421): the file is named "stencils_maxsz<max_st_size>_"
    // followed by the grid file name, and each line is "<size> i0 i1 ..." (the centre first).
    // Both need the maximum stencil size (setMaxStencilSize or generateStencils): without it
    // nothing is written or read, as in the reference. loadStencilsFromFile returns
    // NO_STENCIL_FILES when the file cannot be opened.
    std::string stencilFilename(const std::string& filename) const;
    void writeStencilsToFile(const std::string& filename) const;
    GridLoadErrType loadStencilsFromFile(const std::string& filename);

  protected:
    void checkStencilSize();  // grid_interface.cpp:968-1002: clamps to N
    unsigned int max_st_size = 0;
    double max_st_radius = DBL_MAX;
    unsigned int nb_nodes = 0;
    unsigned int global_num_nodes = 0;
    std::vector<NodeType> node_list;
    std::vector<StencilType> stencil_map;
    std::vector<double> avg_stencil_radii, min_stencil_radii, max_stencil_radii;
};

// GridReader::readNodeList / generate (src/grids/grid_reader.cpp:36-124): whitespace-separated
// "x y z [extra...]" lines; num_extra_cols extra doubles per line are skipped. The constructor
// only records the file; generate() reads it. With max_nodes set (not (unsigned)-1), at most
// that many nodes are read and a file with fewer is an error (print + exit), as in the reference.
class GridReader : public Grid {
  public:
    GridReader(const std::string& filename, int num_extra_cols = 0, unsigned int max_nodes = (unsigned)-1)
        : filename_(filename), extra_(num_extra_cols), n_nodes_(max_nodes) {}
    virtual void generate();
    int readNodeList(int expect_num_extra_dbls_per_line);

  private:
    std::string filename_;
    int extra_;
    unsigned int n_nodes_;
};

// Single-rank stand-in for the reference's Domain (src/grids/domain.h:48-80): every node is in
// Q (owned) and no halo exists. Multi-GPU decomposition is designed but not built.
class Domain : public Grid {
  public:
    Domain() {}
    explicit Domain(const Grid& g) : Grid(g) {}
    unsigned int g2l(unsigned int i) const { return i; }
    unsigned int l2g(unsigned int i) const { return i; }
};

// Single-rank stand-in for the reference's MPI Communicator (src/utils/comm/communicator.h:9-52).
class Communicator {
  public:
    int getRank() const { return 0; }
    int getSize() const { return 1; }
    bool isMaster() const { return true; }
    void barrier() const {}
};

}  // namespace rbfx
