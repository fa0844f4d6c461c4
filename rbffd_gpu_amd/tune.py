"""Autotuner for the fused ELL kernels (SpMM and RK4 stage).

Sweeps the launch shapes the library offers (fields per thread, block size, gather batch),
times each with HIP events on the context stream, installs the fastest with rbfx_tune_set
and keeps a small JSON table (keyed by rows, width, operators, fields, precisions) so a later
run can reuse the choice without sweeping.
"""
import json
import os

from .lib import RK4, EllOperators

KIND_SPMM, KIND_RK4 = 0, 1
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_table.json")


def candidates(n_fields):
    for fpt in (1, 2, 4, 8):
        if n_fields % fpt:
            continue
        for block in (64, 128, 256):
            if n_fields // fpt > block:
                continue
            for batch in (4, 8):
                yield fpt, block, batch


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
        ell.tune_set(kind, n_fields, field_type, accum, e["fpt"], e["block"], e["batch"])
        return e
    return None


def tune_rk4(ctx, rk: RK4, ell: EllOperators, field_type, accum, dt=1e-6, reps=3, table=None, log=None):
    """Times one RK4 step (4 eager stage launches) per shape; installs and returns the best."""
    F = rk.n_fields
    results = []
    for fpt, block, batch in candidates(F):
        ell.tune_set(KIND_RK4, F, field_type, accum, fpt, block, batch)
        for s in (1, 2, 3, 4):
            rk.stage(s, dt)  # warm
        ctx.timer_start()
        for _ in range(reps):
            for s in (1, 2, 3, 4):
                rk.stage(s, dt)
        ms = ctx.timer_stop() / reps
        results.append((ms, fpt, block, batch))
        if log:
            log(f"  rk4 shape fpt={fpt} block={block} batch={batch}: {ms * 1e3:.1f} us/step")
    ms, fpt, block, batch = min(results)
    ell.tune_set(KIND_RK4, F, field_type, accum, fpt, block, batch)
    entry = {"fpt": fpt, "block": block, "batch": batch, "us_per_step": round(ms * 1e3, 2)}
    if table is not None:
        table[_key(ell, KIND_RK4, F, field_type, accum)] = entry
    return entry, sorted(results)


def tune_spmm(ctx, ell: EllOperators, n_fields, d_u, d_out, field_type, accum, reps=5, table=None, log=None):
    results = []
    for fpt, block, batch in candidates(n_fields):
        ell.tune_set(KIND_SPMM, n_fields, field_type, accum, fpt, block, batch)
        ell.spmm(n_fields, d_u, d_out, field_type, accum)
        ctx.timer_start()
        for _ in range(reps):
            ell.spmm(n_fields, d_u, d_out, field_type, accum)
        ms = ctx.timer_stop() / reps
        results.append((ms, fpt, block, batch))
        if log:
            log(f"  spmm shape fpt={fpt} block={block} batch={batch}: {ms * 1e3:.1f} us")
    ms, fpt, block, batch = min(results)
    ell.tune_set(KIND_SPMM, n_fields, field_type, accum, fpt, block, batch)
    entry = {"fpt": fpt, "block": block, "batch": batch, "us": round(ms * 1e3, 2)}
    if table is not None:
        table[_key(ell, KIND_SPMM, n_fields, field_type, accum)] = entry
    return entry, sorted(results)
