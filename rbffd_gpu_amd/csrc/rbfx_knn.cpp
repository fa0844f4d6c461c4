// Host side of stencil generation: seeded node sets and kd-tree k-nearest-neighbour stencils.
//
// The kd-tree follows the reference's algorithm step for step so that stencil index lists
// match it bit for bit, ties included:
//   * tree build: per-dimension heap sort of coordinates (external/kdtree/KDTree.h:179-202,
//     MinHeap push/pop/heapsort external/kdtree/MyHeaps.h:268-455), then recursive median
//     split with the median going left, iMedian = floor((numel-1)/2), children visited in
//     L-then-R order and a stable re-partition of every dimension's sorted list
//     (KDTree.h:267-330);
//   * query: Friedman-Bentley-Finkel search (KDTree.h:459-505) with a bounded max-heap whose
//     replacement rule is "top > d" (KDTree.h:470) and bounds test KDTree.h:566-590; results
//     are popped from the heap and reversed (KDTree.h:407-433);
//   * distance: d = sum over d=0,1,2 of (a_d-b_d)^2 in that order (KDTree.h:46-52), built
//     with -ffp-contract=off so no fused multiply-add changes the rounding;
//   * stencil: row = the k returned ids, no radius trimming (grid_interface.cpp:688-725),
//     k clamped to N (grid_interface.cpp:968-1002).
// The heaps are re-implemented here (the ordering of equal keys depends on their exact
// sift rules, which is why those rules are kept). The tree itself is stored flat with
// in-place stable partitions instead of the reference's per-node vector copies; that changes
// memory use, not the resulting tree. Queries run in parallel (OpenMP), each with its own heap.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "rbfx_host_internal.h"

namespace {

struct KeyIdx {
    double key;
    int idx;
};

// Min-heap with the reference's sift rules: push appends then swaps with the parent while
// parent.key > cur.key; pop moves the tail to the root and sifts down choosing the strictly
// smaller child (left tested first).
class MinHeapD {
  public:
    explicit MinHeapD(size_t cap) { h_.reserve(cap); }
    void push(double key, int idx) {
        h_.push_back({key, idx});
        size_t cur = h_.size() - 1;
        while (cur > 0) {
            size_t par = (cur - 1) >> 1;
            if (h_[par].key > h_[cur].key) {
                std::swap(h_[par], h_[cur]);
                cur = par;
            } else {
                break;
            }
        }
    }
    bool empty() const { return h_.empty(); }
    KeyIdx pop() {
        KeyIdx top = h_[0];
        h_[0] = h_.back();
        h_.pop_back();
        size_t cur = 0, n = h_.size();
        for (;;) {
            size_t l = 2 * cur + 1, r = 2 * cur + 2, s;
            s = (l < n && h_[l].key < h_[cur].key) ? l : cur;
            if (r < n && h_[r].key < h_[s].key) s = r;
            if (s == cur) break;
            std::swap(h_[cur], h_[s]);
            cur = s;
        }
        return top;
    }

  private:
    std::vector<KeyIdx> h_;
};

// Bounded max-heap for the k best candidates, same sift rules mirrored (strict comparisons).
class MaxHeapD {
  public:
    void reset(int cap) {
        n_ = 0;
        if ((int)h_.size() < cap + 1) h_.resize(cap + 1);
    }
    int size() const { return n_; }
    const KeyIdx& top() const { return h_[0]; }
    void push(double key, int idx) {
        int cur = n_++;
        h_[cur] = {key, idx};
        while (cur > 0) {
            int par = (cur - 1) >> 1;
            if (h_[par].key < h_[cur].key) {
                std::swap(h_[par], h_[cur]);
                cur = par;
            } else {
                break;
            }
        }
    }
    void pop() {
        h_[0] = h_[n_ - 1];
        n_--;
        int cur = 0;
        for (;;) {
            int l = 2 * cur + 1, r = 2 * cur + 2, g;
            g = (l < n_ && h_[l].key > h_[cur].key) ? l : cur;
            if (r < n_ && h_[r].key > h_[g].key) g = r;
            if (g == cur) break;
            std::swap(h_[cur], h_[g]);
            cur = g;
        }
    }

  private:
    std::vector<KeyIdx> h_;
    int n_ = 0;
};

struct TreeNode {
    double key;
    int left, right;  // -1 when absent
    int pidx;         // >= 0 for a leaf
};

class KdTree3 {
  public:
    KdTree3(const double* xyz, int n) : xyz_(xyz), n_(n) {
        // sorter[d]: point ids in ascending coordinate d, by heap sort
        for (int d = 0; d < 3; d++) {
            sorter_[d].resize(n);
            MinHeapD heap(n);
            for (int p = 0; p < n; p++) heap.push(xyz[3 * p + d], p);
            int i = 0;
            while (!heap.empty()) sorter_[d][i++] = heap.pop().idx;
        }
        side_.assign(n, 0);
        tmp_.resize(n);
        nodes_.reserve(2 * (size_t)n);
        build(0, n, 0);
        for (int d = 0; d < 3; d++) std::vector<int>().swap(sorter_[d]);
        std::vector<int>().swap(tmp_);
        std::vector<char>().swap(side_);
    }

    // k nearest of point q (ids written closest first).
    void knn(const double* q, int k, MaxHeapD& pq, int* out) const {
        double bmin[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        double bmax[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
        pq.reset(k);
        search(q, k, 0, 0, bmin, bmax, pq);
        int cnt = pq.size();
        // pop all: farthest first; keep them and reverse
        for (int i = cnt - 1; i >= 0; i--) {
            out[i] = pq.top().idx;
            pq.pop();
        }
    }

  private:
    int build(int lo, int hi, int dim) {
        int numel = hi - lo;
        int node_idx = (int)nodes_.size();
        nodes_.push_back({0.0, -1, -1, -1});
        if (numel == 1) {
            nodes_[node_idx].pidx = sorter_[dim][lo];
            nodes_[node_idx].key = 0.0;
            return node_idx;
        }
        int imed = (int)std::floor((numel - 1) / 2.0);
        int pmed = sorter_[dim][lo + imed];
        int nl = imed + 1;
        for (int i = lo; i < hi; i++) side_[sorter_[dim][i]] = (i - lo <= imed) ? 'l' : 'r';
        for (int d = 0; d < 3; d++) {
            // stable partition of sorter_[d][lo,hi) into l then r
            int il = lo, ir = 0;
            for (int i = lo; i < hi; i++) {
                int p = sorter_[d][i];
                if (side_[p] == 'l') sorter_[d][il++] = p;
                else tmp_[ir++] = p;
            }
            std::memcpy(&sorter_[d][lo + nl], tmp_.data(), sizeof(int) * (size_t)ir);
        }
        double key = xyz_[3 * pmed + dim];
        int l = build(lo, lo + nl, (dim + 1) % 3);
        int r = build(lo + nl, hi, (dim + 1) % 3);
        nodes_[node_idx].key = key;
        nodes_[node_idx].left = l;
        nodes_[node_idx].right = r;
        nodes_[node_idx].pidx = -1;
        return node_idx;
    }

    static bool overlap(const double* q, int k, const double* bmin, const double* bmax, const MaxHeapD& pq) {
        if (pq.size() < k) return true;
        double sum = 0.0;
        double best = pq.top().key;
        for (int d = 0; d < 3; d++) {
            if (q[d] < bmin[d]) {
                sum += (q[d] - bmin[d]) * (q[d] - bmin[d]);
                if (sum > best) return false;
            } else if (q[d] > bmax[d]) {
                sum += (q[d] - bmax[d]) * (q[d] - bmax[d]);
                if (sum > best) return false;
            }
        }
        return true;
    }

    void search(const double* q, int k, int node_idx, int dim, double* bmin, double* bmax, MaxHeapD& pq) const {
        const TreeNode& nd = nodes_[node_idx];
        if (nd.pidx >= 0) {
            const double* p = xyz_ + 3 * (size_t)nd.pidx;
            double dist = 0.0;
            for (int i = 0; i < 3; i++) dist += (q[i] - p[i]) * (q[i] - p[i]);
            if (pq.size() == k && pq.top().key > dist) {
                pq.pop();
                pq.push(dist, nd.pidx);
            } else if (pq.size() < k) {
                pq.push(dist, nd.pidx);
            }
            return;
        }
        int nd1 = (dim + 1) % 3;
        double tmp;
        if (q[dim] <= nd.key) {
            tmp = bmax[dim]; bmax[dim] = nd.key;
            search(q, k, nd.left, nd1, bmin, bmax, pq);
            bmax[dim] = tmp;
        } else {
            tmp = bmin[dim]; bmin[dim] = nd.key;
            search(q, k, nd.right, nd1, bmin, bmax, pq);
            bmin[dim] = tmp;
        }
        if (q[dim] <= nd.key) {
            tmp = bmin[dim]; bmin[dim] = nd.key;
            if (overlap(q, k, bmin, bmax, pq)) search(q, k, nd.right, nd1, bmin, bmax, pq);
            bmin[dim] = tmp;
        } else {
            tmp = bmax[dim]; bmax[dim] = nd.key;
            if (overlap(q, k, bmin, bmax, pq)) search(q, k, nd.left, nd1, bmin, bmax, pq);
            bmax[dim] = tmp;
        }
    }

    const double* xyz_;
    int n_;
    std::vector<int> sorter_[3];
    std::vector<char> side_;
    std::vector<int> tmp_;
    std::vector<TreeNode> nodes_;
};

// SplitMix64 (Steele, Lea, Flood 2014): a counter-based 64-bit mixer.
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// uniform double in [-1, 1) from the top 53 bits
inline double unit_sym(uint64_t r) { return (double)(r >> 11) * (2.0 / 9007199254740992.0) - 1.0; }

}  // namespace

// Uniform grid of cubic cells for the GPU k-nearest-neighbour search (rbfx_knn_gpu): bounding
// box lo..; cell size h halved from the largest extent until a non-empty cell holds about
// (k+1)/6 nodes (at least 2) or the dense table would pass 2^25 cells; dims g; nodes sorted
// by cell (cell index (z * gy + y) * gx + x, cell c = sorted nodes cstart[c] .. cstart[c+1],
// coordinates pts, node ids pid). The order of the nodes inside a cell is not fixed (parallel
// scatter); the search result does not depend on it. Returns RBFX_E_INVALID for a
// non-finite coordinate.
int rbfx_knn_grid_bin(int32_t n, const double* xyz, int32_t k, double lo[3], double* h_out, int32_t g_out[3],
                      std::vector<int32_t>& cstart, std::vector<double>& pts, std::vector<int32_t>& pid) {
    double l0 = DBL_MAX, l1 = DBL_MAX, l2 = DBL_MAX, u0 = -DBL_MAX, u1 = -DBL_MAX, u2 = -DBL_MAX;
    int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(min : l0, l1, l2) reduction(max : u0, u1, u2) reduction(+ : bad)
#endif
    for (int32_t i = 0; i < n; i++) {
        const double* p = xyz + 3 * (size_t)i;
        bad += !(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]));
        l0 = std::min(l0, p[0]); l1 = std::min(l1, p[1]); l2 = std::min(l2, p[2]);
        u0 = std::max(u0, p[0]); u1 = std::max(u1, p[1]); u2 = std::max(u2, p[2]);
    }
    if (bad) return RBFX_E_INVALID;
    lo[0] = l0; lo[1] = l1; lo[2] = l2;
    const double hi[3] = {u0, u1, u2};
    auto cell_of = [&](const double* p, const int* gg, double inv) {
        size_t c = 0;
        for (int d = 2; d >= 0; d--) {
            int cd = (int)((p[d] - lo[d]) * inv);
            cd = cd < 0 ? 0 : (cd >= gg[d] ? gg[d] - 1 : cd);
            c = c * gg[d] + cd;
        }
        return c;
    };
    double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(ext > 0.0)) ext = 1.0;
    const double target = std::max(2.0, (k + 1) / 6.0);
    int g[3] = {1, 1, 1};
    double h = ext;
    {
        std::vector<uint8_t> seen;
        for (int G = 4; G <= 4096; G *= 2) {
            const double hh = ext * (1.0 + 1e-12) / G;
            int gg[3];
            for (int d = 0; d < 3; d++) gg[d] = std::max(1, std::min(G, (int)std::ceil((hi[d] - lo[d]) / hh)));
            const size_t total = (size_t)gg[0] * gg[1] * gg[2];
            if (total > ((size_t)1 << 25)) break;
            seen.assign(total, 0);
            const double inv = 1.0 / hh;
#ifdef _OPENMP
#pragma omp parallel for
#endif
            for (int32_t i = 0; i < n; i++)
                __atomic_store_n(&seen[cell_of(xyz + 3 * (size_t)i, gg, inv)], (uint8_t)1, __ATOMIC_RELAXED);
            long long occupied = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : occupied)
#endif
            for (long long c = 0; c < (long long)total; c++) occupied += seen[(size_t)c];
            h = hh;
            std::memcpy(g, gg, sizeof(g));
            if ((double)n / (double)occupied <= target) break;
        }
    }
    const double inv_h = 1.0 / h;
    const size_t n_cells = (size_t)g[0] * g[1] * g[2];
    std::vector<int32_t> cell((size_t)n);
    cstart.assign(n_cells + 1, 0);
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int32_t i = 0; i < n; i++) {
        cell[(size_t)i] = (int32_t)cell_of(xyz + 3 * (size_t)i, g, inv_h);
        __atomic_fetch_add(&cstart[(size_t)cell[(size_t)i] + 1], 1, __ATOMIC_RELAXED);
    }
    // inclusive scan of cstart[1..n_cells]: chunk sums, then chunk offsets
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<long long> part((size_t)nt + 1, 0);
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int t = 0, nth = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nth = omp_get_num_threads();  // <= nt
#endif
        const size_t chunk = (n_cells + nth - 1) / nth;
        const size_t a = std::min(n_cells + 1, 1 + (size_t)t * chunk), b = std::min(n_cells + 1, a + chunk);
        long long run = 0;
        for (size_t c = a; c < b; c++) run += cstart[c], cstart[c] = (int32_t)run;
        part[(size_t)t + 1] = run;
#ifdef _OPENMP
#pragma omp barrier
#pragma omp single
#endif
        for (int q = 0; q < nth; q++) part[(size_t)q + 1] += part[(size_t)q];
        for (size_t c = a; c < b; c++) cstart[c] += (int32_t)part[(size_t)t];
    }
    std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
    pts.resize(3 * (size_t)n);
    pid.resize((size_t)n);
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int32_t i = 0; i < n; i++) {
        const int32_t s = __atomic_fetch_add(&fill[(size_t)cell[(size_t)i]], 1, __ATOMIC_RELAXED);
        std::memcpy(&pts[3 * (size_t)s], &xyz[3 * (size_t)i], 3 * sizeof(double));
        pid[(size_t)s] = i;
    }
    *h_out = h;
    std::memcpy(g_out, g, sizeof(g));
    return RBFX_OK;
}

int rbfx_build_local_tables(const int32_t* idx_layout, int n_rows, int n_pad, int width, int tile_layout, int lt,
                            int cap, std::vector<int32_t>& pos, std::vector<int32_t>& off, std::vector<int32_t>& nodes,
                            std::vector<uint16_t>& loc_idx) {
    auto entry = [&](int p, int j) {  // layout offset of (position p, slot j)
        return ((size_t)(p / tile_layout) * (size_t)width + (size_t)j) * tile_layout + (size_t)(p % tile_layout);
    };
    // greedy tiles inside independent chunks of positions (chunks run in parallel)
    const int chunk = 64 * lt;
    const int n_chunks = (n_rows + chunk - 1) / chunk;
    std::vector<std::vector<int32_t>> c_pos((size_t)n_chunks);                 // tile starts per chunk
    std::vector<std::vector<std::vector<int32_t>>> c_nodes((size_t)n_chunks);  // node list per tile
    loc_idx.assign((size_t)n_pad * width, 0);
    int too_big = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : too_big)
#endif
    for (int c = 0; c < n_chunks; c++) {
        const int p_lo = c * chunk, p_hi = std::min(n_rows, (c + 1) * chunk);
        std::vector<int32_t> u, merged, row((size_t)width);
        int p = p_lo;
        while (p < p_hi) {
            const int start = p;
            u.clear();
            while (p < p_hi && p - start < lt) {
                for (int j = 0; j < width; j++) row[(size_t)j] = idx_layout[entry(p, j)];
                std::sort(row.begin(), row.end());
                merged.clear();
                std::set_union(u.begin(), u.end(), row.begin(), row.end(), std::back_inserter(merged));
                if (cap > 0 && p > start && (int)merged.size() > cap) break;
                u.swap(merged);
                p++;
            }
            if (u.size() > 65535) too_big++;
            for (int q = start; q < p; q++)
                for (int j = 0; j < width; j++) {
                    size_t e = entry(q, j);
                    loc_idx[e] = (uint16_t)(std::lower_bound(u.begin(), u.end(), idx_layout[e]) - u.begin());
                }
            c_pos[(size_t)c].push_back(start);
            c_nodes[(size_t)c].push_back(u);
        }
    }
    if (too_big) return -1;
    pos.clear();
    off.assign(1, 0);
    nodes.clear();
    int k_max = 0;
    for (int c = 0; c < n_chunks; c++)
        for (size_t t = 0; t < c_pos[(size_t)c].size(); t++) {
            const std::vector<int32_t>& u = c_nodes[(size_t)c][t];
            pos.push_back(c_pos[(size_t)c][t]);
            nodes.insert(nodes.end(), u.begin(), u.end());
            off.push_back((int32_t)nodes.size());
            k_max = std::max(k_max, (int)u.size());
        }
    pos.push_back(n_rows);
    return k_max;
}

extern "C" {

int rbfx_knn_kdtree(int32_t n_nodes, const double* xyz, int32_t k, int32_t* out_idx, int32_t n_threads) {
    if (n_nodes <= 0 || !xyz || !out_idx || k <= 0) return RBFX_E_INVALID;
    if (k > n_nodes) return RBFX_E_INVALID;  // caller clamps like Grid::checkStencilSize
    KdTree3 tree(xyz, n_nodes);
#ifdef _OPENMP
    int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
    {
        MaxHeapD pq;
#ifdef _OPENMP
#pragma omp for schedule(static, 1024)
#endif
        for (int32_t i = 0; i < n_nodes; i++) tree.knn(xyz + 3 * (size_t)i, k, pq, out_idx + (size_t)i * k);
    }
    (void)n_threads;
    return RBFX_OK;
}

// kd-tree answers for the listed query rows only. compact = 0: out_idx keeps the n_nodes x k
// layout (other rows are left untouched); compact = 1: the r-th listed row goes to
// out_idx[r * k ..]. Used for the rows that the GPU search hands back.
int rbfx_knn_kdtree_rows(int32_t n_nodes, const double* xyz, int32_t k, const int32_t* rows, int32_t n_rows,
                         int32_t* out_idx, int32_t compact) {
    if (n_rows <= 0) return RBFX_OK;
    KdTree3 tree(xyz, n_nodes);
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        MaxHeapD pq;
#ifdef _OPENMP
#pragma omp for schedule(static, 64)
#endif
        for (int32_t r = 0; r < n_rows; r++) {
            const int32_t i = rows[r];
            tree.knn(xyz + 3 * (size_t)i, k, pq, out_idx + (size_t)(compact ? r : i) * k);
        }
    }
    return RBFX_OK;
}

// Icosahedral node set: the 12 vertices of an icosahedron, each face split into four
// `level` times with every new edge midpoint projected onto the unit sphere;
// N = 10 * 4^level + 2 (level 7: 163842). Vertex order: the 12 corners, then midpoints in
// creation order. out_n receives N; xyz may be NULL to query the count.
int rbfx_nodes_icosahedral(int32_t level, double* xyz, int32_t* out_n) {
    if (level < 0 || level > 10 || !out_n) return RBFX_E_INVALID;
    int64_t n_total = 10 * (int64_t(1) << (2 * level)) + 2;
    *out_n = (int32_t)n_total;
    if (!xyz) return RBFX_OK;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<double> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                             0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    for (int i = 0; i < 12; i++) {
        double r = std::sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        for (int d = 0; d < 3; d++) v[3 * i + d] /= r;
    }
    std::vector<int> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                          3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    v.reserve(3 * (size_t)n_total);
    for (int l = 0; l < level; l++) {
        std::vector<int> nf;
        nf.reserve(f.size() * 4);
        std::unordered_map<uint64_t, int> mid;
        mid.reserve(f.size());
        auto midpoint = [&](int a, int b) {
            uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint64_t)std::max(a, b);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double p[3];
            for (int d = 0; d < 3; d++) p[d] = 0.5 * (v[3 * (size_t)a + d] + v[3 * (size_t)b + d]);
            double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            int id = (int)(v.size() / 3);
            for (int d = 0; d < 3; d++) v.push_back(p[d] / r);
            mid.emplace(key, id);
            return id;
        };
        for (size_t q = 0; q < f.size(); q += 3) {
            int a = f[q], b = f[q + 1], c = f[q + 2];
            int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            int tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), tri, tri + 12);
        }
        f.swap(nf);
    }
    if ((int64_t)(v.size() / 3) != n_total) return RBFX_E_INVALID;
    std::memcpy(xyz, v.data(), v.size() * sizeof(double));
    return RBFX_OK;
}

// Processing order along a 3-D Morton (Z-order) curve over the nodes' bounding box, 21 bits
// per axis: out_order[p] = node visited p-th, ties broken by node index. Nodes that are close
// in space end up close in the order, so rows processed together share stencil neighbours.
int rbfx_order_morton(int32_t n_nodes, const double* xyz, int32_t* out_order) {
    if (n_nodes <= 0 || !xyz || !out_order) return RBFX_E_INVALID;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int32_t i = 0; i < n_nodes; i++)
        for (int d = 0; d < 3; d++) {
            double c = xyz[3 * (size_t)i + d];
            if (!std::isfinite(c)) return RBFX_E_INVALID;
            lo[d] = std::min(lo[d], c);
            hi[d] = std::max(hi[d], c);
        }
    auto spread = [](uint64_t v) {  // 21 bits -> every third bit of 63
        v &= 0x1fffffull;
        v = (v | v << 32) & 0x1f00000000ffffull;
        v = (v | v << 16) & 0x1f0000ff0000ffull;
        v = (v | v << 8) & 0x100f00f00f00f00full;
        v = (v | v << 4) & 0x10c30c30c30c30c3ull;
        v = (v | v << 2) & 0x1249249249249249ull;
        return v;
    };
    std::vector<std::pair<uint64_t, int32_t>> keys((size_t)n_nodes);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t i = 0; i < n_nodes; i++) {
        uint64_t k = 0;
        for (int d = 0; d < 3; d++) {
            double span = hi[d] - lo[d];
            double t = span > 0 ? (xyz[3 * (size_t)i + d] - lo[d]) / span : 0.0;
            uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, std::floor(t * 2097152.0)));
            k |= spread(q) << d;
        }
        keys[(size_t)i] = {k, i};
    }
    std::sort(keys.begin(), keys.end());
    for (int32_t p = 0; p < n_nodes; p++) out_order[p] = keys[(size_t)p].second;
    return RBFX_OK;
}

int rbfx_nodes_fibonacci(int32_t n_nodes, double jitter, uint64_t seed, double* xyz) {
    if (n_nodes <= 0 || !xyz) return RBFX_E_INVALID;
    const double golden = M_PI * (1.0 + std::sqrt(5.0));
    for (int32_t i = 0; i < n_nodes; i++) {
        double t = (double)i + 0.5;
        double z = 1.0 - 2.0 * t / (double)n_nodes;
        double s = std::sqrt(std::max(0.0, 1.0 - z * z));
        double ph = golden * t;
        double p[3] = {s * std::cos(ph), s * std::sin(ph), z};
        for (int d = 0; d < 3; d++) {
            double j = 0.0;
            if (jitter != 0.0) j = jitter * unit_sym(splitmix64(seed * 0x100000001B3ull + 3ull * (uint64_t)i + (uint64_t)d));
            xyz[3 * (size_t)i + d] = p[d] + j;
        }
    }
    return RBFX_OK;
}

}  // extern "C"
