"""Autotuner for the fused ELL kernels (SpMM and RK4 stage).

Sweeps the launch shapes the library offers (fields per thread, block size, gather batch
and, for the RK4 stage kernel, whether the lanes of a row share the index/weight loads),
times each with HIP events on the context stream, installs the fastest with rbfx_tune_set
and keeps a small JSON table (keyed by rows, width, operators, fields, precisions) so a later
run can reuse the choice without sweeping.
"""
import json
import os

from .lib import RK4, EllOperators

KIND_SPMM, KIND_RK4 = 0, 1
LDS_BYTES = 160 * 1024  # LDS per workgroup on gfx950
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_table.json")


def _pipe_ok(n_ops, fpt, G, share):
    """Pipelined shared-load cores the library instantiates (share 2/3 = prefetch depth)."""
    return n_ops >= 3 and ((fpt == 4 and G == 4) or (fpt == 8 and G == 2) or
                           (share == 2 and ((fpt == 4 and G == 2) or (fpt == 2 and G == 4))))


def candidates(n_fields, kind=KIND_SPMM, n_ops=1, local=None, field_bytes=8, fp32_accum=False):
    """(fpt, block, batch, share) shapes the library accepts for n_fields fields.
    local = (tile_rows, k_max) of an operator set with LDS-staging tables: adds the staged
    shapes (share 4; the block is the tile) when a tile's records fit in LDS_BYTES, and with
    FP32 fields and accumulation (fp32_accum) the staged shapes whose multiply-adds run on
    the matrix cores (share 5, 16 fields on 4 lanes)."""
    if local and local[0] > 0 and n_ops >= 3 and (n_fields * field_bytes) % 16 == 0 \
            and local[1] * n_fields * field_bytes <= LDS_BYTES:
        for fpt, G in ((4, 4), (8, 2)):
            if n_fields == fpt * G and local[0] * G <= 512:
                for batch in (4, 8):
                    yield fpt, local[0] * G, batch, 4
                    if fp32_accum and field_bytes == 4 and fpt == 4:
                        yield fpt, local[0] * G, batch, 5
    for fpt in (1, 2, 4, 8):
        if n_fields % fpt:
            continue
        G = n_fields // fpt
        for block in (64, 128, 256):
            if G > block:
                continue
            for batch in ((4, 8, 16) if (kind == KIND_SPMM and fpt == 1) else (4, 8)):
                yield fpt, block, batch, 0
                if kind == KIND_RK4 and fpt > 1 and G in (2, 4) and batch % G == 0:
                    yield fpt, block, batch, 1
                    for share in (2, 3):
                        if _pipe_ok(n_ops, fpt, G, share):
                            yield fpt, block, batch, share


def _key(ell, kind, n_fields, field_type, accum):
    return f"k{kind}_r{ell.n_rows}_w{ell.width}_o{ell.n_ops}_f{n_fields}_s{ell.storage}_t{field_type}_a{accum}"


def load_table(path=DEFAULT_TABLE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_table(table, path=DEFAULT_TABLE):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def apply_table(ell: EllOperators, kind, n_fields, field_type, accum, table):
    e = table.get(_key(ell, kind, n_fields, field_type, accum))
    if e:
        ell.tune_set(kind, n_fields, field_type, accum, e["fpt"], e["block"], e["batch"], e.get("share", 0))
        return e
    return None


def _confirm(results, time_shape, finalists=4, rounds=3):
    """Re-times the fastest shapes of a first pass in interleaved rounds and ranks them by
    their best time: one noisy first-pass sample then cannot decide the pick."""
    top = sorted(results)[:finalists]
    best = {r[1:]: r[0] for r in top}
    for _ in range(rounds):
        for r in top:
            best[r[1:]] = min(best[r[1:]], time_shape(*r[1:]))
    return sorted((ms, *shape) for shape, ms in best.items()) + sorted(results)[finalists:]


def tune_rk4(ctx, rk: RK4, ell: EllOperators, field_type, accum, dt=1e-6, reps=3, table=None, log=None):
    """Times one RK4 step (4 eager stage launches) per shape, re-times the fastest few in
    interleaved rounds, installs and returns the best."""
    F = rk.n_fields
    local = ell.local_info()
    fp32 = field_type == 1 and accum == 1 and ell.storage != 0

    def time_shape(fpt, block, batch, share, n=reps):
        ell.tune_set(KIND_RK4, F, field_type, accum, fpt, block, batch, share)
        for s in (1, 2, 3, 4):
            rk.stage(s, dt)  # warm
        ctx.timer_start()
        for _ in range(n):
            for s in (1, 2, 3, 4):
                rk.stage(s, dt)
        return ctx.timer_stop() / n

    results = []
    for shape in candidates(F, KIND_RK4, ell.n_ops, local, 8 if field_type == 0 else 4, fp32):
        ms = time_shape(*shape)
        results.append((ms, *shape))
        if log:
            log(f"  rk4 shape fpt={shape[0]} block={shape[1]} batch={shape[2]} share={shape[3]}: {ms * 1e3:.1f} us/step")
    results = _confirm(results, time_shape)
    ms, fpt, block, batch, share = results[0]
    ell.tune_set(KIND_RK4, F, field_type, accum, fpt, block, batch, share)
    entry = {"fpt": fpt, "block": block, "batch": batch, "share": share, "us_per_step": round(ms * 1e3, 2)}
    if table is not None:
        table[_key(ell, KIND_RK4, F, field_type, accum)] = entry
    return entry, results


def tune_spmm(ctx, ell: EllOperators, n_fields, d_u, d_out, field_type, accum, reps=5, table=None, log=None):
    """Times the fused SpMM per shape, re-times the fastest few in interleaved rounds,
    installs and returns the best."""
    local = ell.local_info() if ell.n_ops >= 3 else None
    fp32 = field_type == 1 and accum == 1 and ell.storage != 0

    def time_shape(fpt, block, batch, share, n=reps):
        ell.tune_set(KIND_SPMM, n_fields, field_type, accum, fpt, block, batch, share)
        ell.spmm(n_fields, d_u, d_out, field_type, accum)
        ctx.timer_start()
        for _ in range(n):
            ell.spmm(n_fields, d_u, d_out, field_type, accum)
        return ctx.timer_stop() / n

    results = []
    for shape in candidates(n_fields, KIND_SPMM, ell.n_ops, local, 8 if field_type == 0 else 4, fp32):
        ms = time_shape(*shape)
        results.append((ms, *shape))
        if log:
            log(f"  spmm shape fpt={shape[0]} block={shape[1]} batch={shape[2]} share={shape[3]}: {ms * 1e3:.1f} us")
    results = _confirm(results, time_shape)
    ms, fpt, block, batch, share = results[0]
    ell.tune_set(KIND_SPMM, n_fields, field_type, accum, fpt, block, batch, share)
    entry = {"fpt": fpt, "block": block, "batch": batch, "share": share, "us": round(ms * 1e3, 2)}
    if table is not None:
        table[_key(ell, KIND_SPMM, n_fields, field_type, accum)] = entry
    return entry, results
