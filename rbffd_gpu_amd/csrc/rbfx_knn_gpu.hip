// k-nearest-neighbour stencils on the GPU, bit-identical to the host kd-tree (rbfx_knn_kdtree,
// which follows the reference's external/kdtree/KDTree.h search step for step).
//
// Why an exact search suffices: the kd-tree returns its bounded max-heap's contents in
// ascending distance (KDTree.h:407-433), and its pruning keeps every point whose distance is
// below the k-th best. When the k+1 smallest distances of a query are all distinct, the result
// is therefore the unique list of the k nearest points in ascending distance, whatever order
// the tree visits them in. Only when two of those k+1 distances are equal do the heap's sift
// rules and the visit order decide which point comes first (or which of two equidistant
// points at the k-th place is kept). So:
//   * the GPU runs an exact search over a uniform grid of cubic cells, one query per thread,
//     keeping the k+1 best (distance, id) pairs in LDS in ascending order;
//   * distances are computed exactly as the kd-tree does, (q0-p0)^2 + (q1-p1)^2 + (q2-p2)^2
//     summed left to right with every product and sum rounded on its own (no fused
//     multiply-add), so equal inputs give equal keys on both sides;
//   * a query whose k+1 best distances are not strictly increasing is handed back and
//     recomputed on the host by the kd-tree (rbfx_knn_kdtree_rows).
// The search visits cells in shells of growing Chebyshev radius r around the query's cell;
// every point outside the cells visited so far is at least r * h away, so the search stops
// once the (k+1)-th best distance is below (r h)^2 (with a 1e-9 relative margin for the
// rounding of cell assignment and distances), and at the latest when r covers the grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "rbfx_host_internal.h"
#include "rbfx_internal.h"

namespace {

struct KnnArgs {
    const double* pts;      // points sorted by cell, xyz interleaved
    const int32_t* pid;     // node id of each sorted point
    const int32_t* cstart;  // cell c holds sorted points cstart[c] .. cstart[c+1]
    const double* q;        // query coordinates (the nodes, in node order)
    int32_t* out;           // n x k neighbour ids
    int32_t* flag;          // 1: distance tie among the k+1 best, recompute on the host
    int n, k, gx, gy, gz;
    double lo0, lo1, lo2, inv_h, h;
};

__device__ __forceinline__ int cell_of(double x, double lo, double inv_h, int g) {
    const int c = (int)((x - lo) * inv_h);  // x >= lo
    return c < 0 ? 0 : (c >= g ? g - 1 : c);
}

// One query per thread; the sorted k+1 best live in LDS at [slot * B + thread].
__global__ void __launch_bounds__(64) knn_grid_kernel(KnnArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int B = blockDim.x, t = threadIdx.x, K1 = a.k + 1;
    double* dl = reinterpret_cast<double*>(smem);
    int* il = reinterpret_cast<int*>(dl + (size_t)K1 * B);
    const int i = blockIdx.x * B + t;
    if (i >= a.n) return;  // the kernel has no barriers
    const double q0 = a.q[3 * (size_t)i], q1 = a.q[3 * (size_t)i + 1], q2 = a.q[3 * (size_t)i + 2];
    const int cx = cell_of(q0, a.lo0, a.inv_h, a.gx), cy = cell_of(q1, a.lo1, a.inv_h, a.gy),
              cz = cell_of(q2, a.lo2, a.inv_h, a.gz);
    const int rmax = max(a.gx, max(a.gy, a.gz));  // at rmax every cell has been visited
    int cnt = 0;
    for (int r = 0; r <= rmax; r++) {
        for (int dz = -r; dz <= r; dz++) {
            const int z = cz + dz;
            if (z < 0 || z >= a.gz) continue;
            for (int dy = -r; dy <= r; dy++) {
                const int y = cy + dy;
                if (y < 0 || y >= a.gy) continue;
                // on the shell's faces every dx; inside it only dx = -r and dx = +r
                const int step = (dz == -r || dz == r || dy == -r || dy == r || r == 0) ? 1 : 2 * r;
                for (int dx = -r; dx <= r; dx += step) {
                    const int x = cx + dx;
                    if (x < 0 || x >= a.gx) continue;
                    const size_t c = ((size_t)z * a.gy + y) * a.gx + x;
                    const int s1 = a.cstart[c + 1];
                    for (int s = a.cstart[c]; s < s1; s++) {
                        const double e0 = q0 - a.pts[3 * (size_t)s], e1 = q1 - a.pts[3 * (size_t)s + 1],
                                     e2 = q2 - a.pts[3 * (size_t)s + 2];
                        const double d = dadd_nc(dadd_nc(dmul_nc(e0, e0), dmul_nc(e1, e1)), dmul_nc(e2, e2));
                        int pos;
                        if (cnt < K1) {
                            pos = cnt++;
                        } else if (d < dl[(size_t)(K1 - 1) * B + t]) {
                            pos = K1 - 1;
                        } else {
                            continue;
                        }
                        while (pos > 0 && dl[(size_t)(pos - 1) * B + t] > d) {
                            dl[(size_t)pos * B + t] = dl[(size_t)(pos - 1) * B + t];
                            il[(size_t)pos * B + t] = il[(size_t)(pos - 1) * B + t];
                            pos--;
                        }
                        dl[(size_t)pos * B + t] = d;
                        il[(size_t)pos * B + t] = a.pid[s];
                    }
                }
            }
        }
        if (cnt == K1) {
            const double bound = (double)r * a.h;
            if (dl[(size_t)(K1 - 1) * B + t] < bound * bound * (1.0 - 1e-9)) break;
        }
    }
    int tie = 0;
    for (int j = 0; j < a.k; j++) tie |= dl[(size_t)j * B + t] == dl[(size_t)(j + 1) * B + t];
    a.flag[i] = tie;
    for (int j = 0; j < a.k; j++) a.out[(size_t)i * a.k + j] = il[(size_t)j * B + t];
}

struct DevBufs {
    void* p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    ~DevBufs() {
        for (void* q : p)
            if (q) hipFree(q);
    }
};

// The search with its output either in host memory (out_host) or in device memory (d_out).
int knn_gpu_impl(rbfx_ctx* ctx, int32_t n_nodes, const double* xyz, int32_t k, int32_t* out_host, int32_t* d_out,
                 int32_t* n_host_rows) {
    if (n_host_rows) *n_host_rows = 0;
    RBFX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // host kd-tree answers for `rows`, written to the output wherever it lives
    auto host_rows = [&](const std::vector<int32_t>& rows) -> int {
        if (n_host_rows) *n_host_rows = (int32_t)rows.size();
        if (rows.empty()) return RBFX_OK;
        if (out_host) return rbfx_knn_kdtree_rows(n_nodes, xyz, k, rows.data(), (int32_t)rows.size(), out_host, 0);
        std::vector<int32_t> tmp(rows.size() * (size_t)k);
        int rc = rbfx_knn_kdtree_rows(n_nodes, xyz, k, rows.data(), (int32_t)rows.size(), tmp.data(), 1);
        if (rc != RBFX_OK) return rc;
        for (size_t r = 0; r < rows.size(); r++)
            RBFX_HIP_CHECK(ctx, hipMemcpyAsync(d_out + (size_t)rows[r] * k, tmp.data() + r * k, sizeof(int32_t) * k,
                                               hipMemcpyHostToDevice, st));
        RBFX_HIP_CHECK(ctx, hipStreamSynchronize(st));
        return RBFX_OK;
    };
    // the GPU list holds k+1 entries per query in LDS (64 or 32 queries per block)
    const int B = (size_t)(k + 1) * 64 * 12 <= 65536 ? 64 : 32;
    if (k + 1 > n_nodes || (size_t)(k + 1) * B * 12 > 65536) {
        for (size_t i = 0; i < 3 * (size_t)n_nodes; i++)
            if (!std::isfinite(xyz[i])) return rbfx_fail(ctx, RBFX_E_INVALID, "knn_gpu: non-finite coordinate");
        std::vector<int32_t> rows((size_t)n_nodes);
        for (int32_t i = 0; i < n_nodes; i++) rows[(size_t)i] = i;
        return host_rows(rows);
    }
    double lo[3], h = 1.0;
    int32_t g[3];
    std::vector<int32_t> cstart, pid;
    std::vector<double> pts;
    if (rbfx_knn_grid_bin(n_nodes, xyz, k, lo, &h, g, cstart, pts, pid) != RBFX_OK)
        return rbfx_fail(ctx, RBFX_E_INVALID, "knn_gpu: non-finite coordinate");
    const size_t n_cells = (size_t)g[0] * g[1] * g[2];
    DevBufs db;
    const size_t sz[6] = {sizeof(double) * 3 * (size_t)n_nodes, sizeof(int32_t) * (size_t)n_nodes,
                          sizeof(int32_t) * (n_cells + 1), sizeof(double) * 3 * (size_t)n_nodes,
                          d_out ? 0 : sizeof(int32_t) * (size_t)n_nodes * k, sizeof(int32_t) * (size_t)n_nodes};
    for (int b = 0; b < 6; b++) {
        if (!sz[b]) continue;
        hipError_t e = hipMalloc(&db.p[b], sz[b]);
        if (e != hipSuccess) return rbfx_fail(ctx, RBFX_E_NOMEM, "knn_gpu: %s", hipGetErrorString(e));
    }
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(db.p[0], pts.data(), sz[0], hipMemcpyHostToDevice, st));
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(db.p[1], pid.data(), sz[1], hipMemcpyHostToDevice, st));
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(db.p[2], cstart.data(), sz[2], hipMemcpyHostToDevice, st));
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(db.p[3], xyz, sz[3], hipMemcpyHostToDevice, st));
    KnnArgs a;
    a.pts = (const double*)db.p[0];
    a.pid = (const int32_t*)db.p[1];
    a.cstart = (const int32_t*)db.p[2];
    a.q = (const double*)db.p[3];
    a.out = d_out ? d_out : (int32_t*)db.p[4];
    a.flag = (int32_t*)db.p[5];
    a.n = n_nodes;
    a.k = k;
    a.gx = g[0];
    a.gy = g[1];
    a.gz = g[2];
    a.lo0 = lo[0];
    a.lo1 = lo[1];
    a.lo2 = lo[2];
    a.inv_h = 1.0 / h;
    a.h = h;
    const size_t lds = (size_t)(k + 1) * B * 12;
    knn_grid_kernel<<<(n_nodes + B - 1) / B, B, lds, st>>>(a);
    RBFX_LAUNCH_CHECK(ctx, "knn_grid_kernel");
    std::vector<int32_t> flag((size_t)n_nodes);
    if (out_host)
        RBFX_HIP_CHECK(ctx, hipMemcpyAsync(out_host, db.p[4], sizeof(int32_t) * (size_t)n_nodes * k,
                                           hipMemcpyDeviceToHost, st));
    RBFX_HIP_CHECK(ctx, hipMemcpyAsync(flag.data(), db.p[5], sz[5], hipMemcpyDeviceToHost, st));
    RBFX_HIP_CHECK(ctx, hipStreamSynchronize(st));
    std::vector<int32_t> rows;
    for (int32_t i = 0; i < n_nodes; i++)
        if (flag[(size_t)i]) rows.push_back(i);
    return host_rows(rows);
}

}  // namespace

extern "C" {

// Exact k-nearest-neighbour stencils, bit-identical to rbfx_knn_kdtree (host arrays in and
// out). Replaces the query loop of Grid::generateStencils(ST_KDTREE)
// (grid_interface.cpp:688-725) with a GPU search; rows with distance ties come from the
// kd-tree on the host, and n_host_rows (optional) receives how many rows that was.
int rbfx_knn_gpu(rbfx_ctx* ctx, int32_t n_nodes, const double* xyz, int32_t k, int32_t* out_idx,
                 int32_t* n_host_rows) {
    if (!ctx) return RBFX_E_INVALID;
    if (n_nodes <= 0 || !xyz || !out_idx || k <= 0 || k > n_nodes)
        return rbfx_fail(ctx, RBFX_E_INVALID, "knn_gpu: bad arguments (n_nodes %d, k %d)", n_nodes, k);
    return knn_gpu_impl(ctx, n_nodes, xyz, k, out_idx, nullptr, n_host_rows);
}

// The same lists written to device memory (d_out_idx: n_nodes * k int32 on the context's
// device), so that the weights and operators can be built from them without a round trip
// through host memory. Tie rows are recomputed on the host and uploaded row by row.
int rbfx_knn_gpu_device(rbfx_ctx* ctx, int32_t n_nodes, const double* xyz, int32_t k, int32_t* d_out_idx,
                        int32_t* n_host_rows) {
    if (!ctx) return RBFX_E_INVALID;
    if (n_nodes <= 0 || !xyz || !d_out_idx || k <= 0 || k > n_nodes)
        return rbfx_fail(ctx, RBFX_E_INVALID, "knn_gpu_device: bad arguments (n_nodes %d, k %d)", n_nodes, k);
    return knn_gpu_impl(ctx, n_nodes, xyz, k, nullptr, d_out_idx, n_host_rows);
}

}  // extern "C"
