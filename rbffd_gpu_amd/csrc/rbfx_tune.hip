// Launch-shape selection for the fused ELL kernels (fields per thread, block size).
//
// A tuning table maps (kind, rows bucket, width, n_ops, fields, storage, field type,
// accumulation) to the shape measured fastest by rbfx_tune_spmm / the autotuner in the
// Python layer; the table lives in memory and can be saved/loaded as text so a run can reuse
// an earlier sweep. Without an entry: an RK4 stage of an operator set with LDS staging tables
// (rbfx_ell_build_local) that fit takes the LDS-staged shape (4 lanes per row, 16 fields,
// batch 4; the matrix-core variant for FP32 weights, fields and accumulation); otherwise the
// heuristic picks the largest fields-per-thread of 8/4/2/1 that divides F and keeps at most
// 16 FP64 (32 FP32) accumulators per thread.
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>

#include "rbfx_ell_common.h"

namespace {
using Key = std::tuple<int, int, int, int, int, int, int, int>;
std::map<Key, EllTune> g_table;
std::mutex g_mu;

int rows_bucket(int n) {
    int b = 0;
    while ((1 << (b + 1)) <= n) b++;
    return b;
}

Key make_key(const rbfx_ell* e, int F, int ft, int acc, int kind) {
    return Key(kind, rows_bucket(e->n_rows), e->width, e->n_ops, F, e->storage, ft, acc);
}

bool valid(int F, int fpt, int block) {
    if (fpt <= 0 || F % fpt) return false;
    if (fpt != 1 && fpt != 2 && fpt != 4 && fpt != 8) return false;
    if (block < 64 || block > 256 || block % 64) return false;
    return (F / fpt) <= block;
}

// An LDS-staged shape (share 4/5) holds for the operator set it was tuned on: another set with
// the same key may have no tile tables, other tiles or more nodes per tile than the LDS holds.
bool lds_shape_fits(const rbfx_ell* e, int F, int ft, int acc, int kind, const EllTune& c) {
    if (c.fpt <= 0 || F % c.fpt) return false;
    const int G = F / c.fpt;
    const size_t rec = (size_t)F * (ft == RBFX_F64 ? 8 : 4);
    bool ok = (kind == 0 || kind == 1) && e->lt > 0 && e->n_ops >= 3 && e->lt * G <= 512 && rec % 16 == 0 &&
              (size_t)e->k_max * rec <= (size_t)e->lds_limit;
    if (c.share == 5) ok = ok && c.fpt == 4 && G == 4 && e->storage != RBFX_F64 && ft == RBFX_F32 && acc == RBFX_F32;
    return ok;
}
}  // namespace

int rbfx_ell_pick(rbfx_ell* e, int F, int ft, int acc, int kind, EllTune* t) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_table.find(make_key(e, F, ft, acc, kind));
        if (it != g_table.end()) {
            EllTune c = it->second;
            if (c.share != 4 && c.share != 5) {
                *t = c;
                return RBFX_OK;
            }
            if (lds_shape_fits(e, F, ft, acc, kind, c)) {
                c.block = e->lt * (F / c.fpt);  // the block is this set's tile
                *t = c;
                return RBFX_OK;
            }
            // otherwise: the shape of a set without (fitting) tile tables, below
        }
    }
    if ((kind == 0 || kind == 1) && e->lt > 0 && e->n_ops >= 3 && F == 16 && e->lt * 4 <= 512) {
        const size_t rec = (size_t)F * (ft == RBFX_F64 ? 8 : 4);
        if ((size_t)e->k_max * rec <= (size_t)e->lds_limit) {
            const bool mx = e->storage != RBFX_F64 && ft == RBFX_F32 && acc == RBFX_F32;
            *t = EllTune{4, e->lt * 4, 4, mx ? 5 : 4};
            return RBFX_OK;
        }
    }
    int max_acc = (acc == RBFX_F64) ? 16 : 32;
    int fpt = 1;
    for (int c : {8, 4, 2}) {
        if (F % c == 0 && c * e->n_ops <= max_acc) { fpt = c; break; }
    }
    t->fpt = fpt;
    t->block = 256;
    if (!valid(F, t->fpt, t->block))
        return rbfx_fail(e->ctx, RBFX_E_TOO_LARGE, "no launch shape for %d fields (at most 256*8)", F);
    return RBFX_OK;
}

extern "C" {

// Sets the launch shape used for (ell, n_fields, types, kind); kind 0 = SpMM, 1 = RK4 stage.
int rbfx_tune_set(rbfx_ell* e, int32_t kind, int32_t n_fields, int32_t field_type, int32_t accum, int32_t fpt, int32_t block,
                  int32_t batch, int32_t share) {
    if (!e) return RBFX_E_INVALID;
    if (share == 4 || share == 5) {  // LDS-staged gathers (rbfx_ell_build_local); the block is the tile
        const int G = fpt > 0 ? n_fields / fpt : 0;
        const size_t rec = (size_t)n_fields * (field_type == RBFX_F64 ? 8 : 4);
        bool ok = (kind == 0 || kind == 1) && e->lt > 0 && e->n_ops >= 3 && fpt > 0 && n_fields % fpt == 0 &&
                  ((fpt == 4 && G == 4) || (fpt == 8 && G == 2)) && (batch == 4 || batch == 8) &&
                  e->lt * G <= 512 && rec % 16 == 0 && (size_t)e->k_max * rec <= (size_t)e->lds_limit;
        // matrix-core multiply-adds: FP32 weights, fields and accumulation, 16 fields on 4 lanes
        if (share == 5)
            ok = ok && fpt == 4 && G == 4 && e->storage != RBFX_F64 && field_type == RBFX_F32 && accum == RBFX_F32;
        if (!ok)
            return rbfx_fail(e->ctx, RBFX_E_INVALID,
                             "tune_set: no LDS-staged shape for %d fields, fpt %d, batch %d (tile %d, %d nodes per tile)",
                             n_fields, fpt, batch, e->lt, e->k_max);
        std::lock_guard<std::mutex> lk(g_mu);
        g_table[make_key(e, n_fields, field_type, accum, kind)] = EllTune{fpt, e->lt * G, batch, share};
        return RBFX_OK;
    }
    // batch 16: the SpMV/SpMM kernel with one field per thread only
    const bool batch_ok = batch == 4 || batch == 8 || (batch == 16 && kind == 0 && fpt == 1 && share == 0);
    if (!valid(n_fields, fpt, block) || !batch_ok || share < 0 || share > 3)
        return rbfx_fail(e->ctx, RBFX_E_INVALID, "tune_set: invalid shape fpt %d block %d batch %d share %d for %d fields",
                         fpt, block, batch, share, n_fields);
    if (share) {
        int G = n_fields / fpt;
        if (kind != 1 || (G != 2 && G != 4) || fpt == 1 || batch % G)
            return rbfx_fail(e->ctx, RBFX_E_INVALID,
                             "tune_set: shared loads need the RK4 stage kernel, 2 or 4 lanes per row, fpt > 1 and a "
                             "batch divisible by them (fields %d, fpt %d, batch %d)", n_fields, fpt, batch);
        // pipelined cores (share 2, 3 = prefetch depth) exist for 3-4 operators and these shapes
        const bool pipe_ok = e->n_ops >= 3 && ((fpt == 4 && G == 4) || (fpt == 8 && G == 2) ||
                                               (share == 2 && ((fpt == 4 && G == 2) || (fpt == 2 && G == 4))));
        if (share >= 2 && !pipe_ok)
            return rbfx_fail(e->ctx, RBFX_E_INVALID, "tune_set: no pipelined core for %d operators, fpt %d, %d lanes, depth %d",
                             e->n_ops, fpt, G, share);
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[make_key(e, n_fields, field_type, accum, kind)] = EllTune{fpt, block, batch, share};
    return RBFX_OK;
}

int rbfx_tune_get(rbfx_ell* e, int32_t kind, int32_t n_fields, int32_t field_type, int32_t accum, int32_t* fpt, int32_t* block,
                  int32_t* batch, int32_t* share) {
    if (!e || !fpt || !block || !batch || !share) return RBFX_E_INVALID;
    EllTune t;
    int rc = rbfx_ell_pick(e, n_fields, field_type, accum, kind, &t);
    if (rc) return rc;
    *fpt = t.fpt;
    *block = t.block;
    *batch = t.ub;
    *share = t.share;
    return RBFX_OK;
}

int rbfx_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
    return RBFX_OK;
}

}  // extern "C"
