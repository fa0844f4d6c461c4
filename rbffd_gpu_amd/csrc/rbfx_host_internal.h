// Shared declarations for the host-only translation units of librbfx.
#pragma once
#include <cstdint>
#include <vector>

#include "rbfx.h"

// Local-index tables for LDS-staged gathers, built on the host from an operator set's index
// array in its device layout (tiles of tile_layout rows, slot-major inside a tile, stencil
// width `width`; the first n_rows positions are real). Positions are cut greedily into
// tiles of at most lt consecutive positions whose distinct neighbour nodes number at most
// cap (cap <= 0: no limit; a single row always forms a tile): tile t covers positions
// pos[t] .. pos[t+1]. Per tile, the sorted distinct neighbour nodes (appended to nodes, tile
// t at off[t] .. off[t+1]) and, at every entry of the layout, the neighbour's index in its
// tile's list (loc_idx, same layout; 0 in pad rows). Returns the largest list length, or -1
// when a list would not fit 16-bit local indices.
int rbfx_build_local_tables(const int32_t* idx_layout, int n_rows, int n_pad, int width, int tile_layout, int lt,
                            int cap, std::vector<int32_t>& pos, std::vector<int32_t>& off, std::vector<int32_t>& nodes,
                            std::vector<uint16_t>& loc_idx);

// kd-tree k-nearest-neighbour rows (same answers as rbfx_knn_kdtree) for the query rows
// rows[0..n_rows) only, written at out_idx[row * k ..] (compact: at out_idx[r * k ..] for the
// r-th listed row); k <= n_nodes, rows in range.
extern "C" __attribute__((visibility("hidden"))) int rbfx_knn_kdtree_rows(int32_t n_nodes, const double* xyz, int32_t k, const int32_t* rows,
                                    int32_t n_rows, int32_t* out_idx, int32_t compact);

// Uniform-grid binning of the nodes for the GPU search of rbfx_knn_gpu: bounding box corner
// lo, cell size h, cell counts g, cell start table (size g0*g1*g2 + 1) and the nodes sorted
// by cell (coordinates pts, ids pid). RBFX_E_INVALID for non-finite coordinates.
__attribute__((visibility("hidden"))) int rbfx_knn_grid_bin(int32_t n, const double* xyz, int32_t k, double lo[3],
                                                            double* h_out, int32_t g_out[3],
                                                            std::vector<int32_t>& cstart, std::vector<double>& pts,
                                                            std::vector<int32_t>& pid);
