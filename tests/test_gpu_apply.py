"""GPU parity of derivative application (SpMV / fused SpMM) against the CPU restatement.

Parity metric for FP64 (cancellation-aware, derivative weights sum to zero):
    |y_gpu - y_cpu| / sum_j |w_ij u_j|  <= 1e-12  per row.
FP32-stored weights/fields are compared against the FP64 result with the same
normalisation; the bound is the FP32 rounding of the inputs (a few u32 = 6e-8).
"""
import numpy as np
import pytest

import oracle as O
import rbffd_gpu_amd as R
from conftest import rel_apply_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6


def _absum(idx, w, u):
    """sum_j |w_ij u_idx(ij)| per row (and field)."""
    return np.einsum("ij,ijf->if", np.abs(w), np.abs(u[idx])) if u.ndim == 2 else (np.abs(w) * np.abs(u[idx])).sum(1)


def test_spmv_c1_reference_weights(ctx, c1):
    """Config 1: oracle weights (X,Y,Z,LSFC; deg 2) applied on the GPU, one operator at a time."""
    print("case: spmv c1 N=4096 n=31", flush=True)
    xyz, st = c1["xyz"], c1["st"]
    off, idx = R.csr_from_fixed(st)
    mask = R.X | R.Y | R.Z | R.LSFC
    w, _, ns = O.weights(off, idx, xyz, 3.5, mask, dim=3, poly_deg=2)
    assert ns == 0
    x, y, z = xyz.T
    f = np.exp(x) * np.sin(2 * y) + z ** 3
    d_f = ctx.to_device(f)
    d_out = ctx.alloc(4096, np.float64)
    for t in range(4):
        ref = O.apply(off, idx, w[t], f)
        ell = R.EllOperators(ctx, st, [w[t]], storage=R.F64)
        ell.spmv(0, d_f, d_out)
        got = d_out.download()
        ell.close()
        err = rel_apply_err(got, ref, _absum(st, w[t].reshape(st.shape), f))
        print(f"  type {t}: normalised err {err:.2e}")
        assert err <= TOL64


def test_spmv_interleaved_operator(ctx, c1):
    """Single operator out of a 4-operator interleaved set (strided weights)."""
    print("case: spmv strided c1", flush=True)
    xyz, st = c1["xyz"], c1["st"]
    rng = np.random.default_rng(3)
    ws = [rng.standard_normal(st.shape) for _ in range(4)]
    u = rng.standard_normal(4096)
    ell = R.EllOperators(ctx, st, ws, storage=R.F64)
    d_u = ctx.to_device(u)
    d_o = ctx.alloc(4096, np.float64)
    for o in range(4):
        ell.spmv(o, d_u, d_o)
        off, idx = R.csr_from_fixed(st)
        ref = O.apply(off, idx, ws[o].reshape(-1), u)
        assert rel_apply_err(d_o.download(), ref, _absum(st, ws[o], u)) <= TOL64
    ell.close()


def test_apply_exact_integer_kat(ctx):
    """Small-integer weights and fields: every partial sum is exact in FP64 and FP32."""
    print("case: integer KAT", flush=True)
    rng = np.random.default_rng(5)
    n_rows, width = 1000, 17
    idx = rng.integers(0, n_rows, size=(n_rows, width)).astype(np.int32)
    idx[:, 0] = np.arange(n_rows)
    w = rng.integers(-8, 9, size=(n_rows, width)).astype(np.float64)
    u = rng.integers(-16, 17, size=n_rows).astype(np.float64)
    exact = (w * u[idx]).sum(1)
    for storage, ft, acc in ((R.F64, R.F64, R.F64), (R.F32, R.F32, R.F64), (R.F32, R.F32, R.F32)):
        ell = R.EllOperators(ctx, idx, [w], storage=storage)
        dt = np.float64 if ft == R.F64 else np.float32
        d_u = ctx.to_device(u.astype(dt))
        d_o = ctx.alloc(n_rows, dt)
        ell.spmv(0, d_u, d_o, ft, acc)
        assert np.array_equal(d_o.download().astype(np.float64), exact), (storage, ft, acc)
        ell.close()


@pytest.mark.parametrize("n_ops,F", [(1, 1), (3, 1), (4, 16), (4, 8), (2, 6), (4, 3)])
def test_spmm_fused(ctx, n_ops, F):
    """Fused multi-operator multi-field SpMM, FP64 and FP32-stored, vs the CPU restatement."""
    print(f"case: spmm n_ops={n_ops} F={F}", flush=True)
    rng = np.random.default_rng(11 + n_ops * 7 + F)
    n_rows, width = 5000, 50
    idx = rng.integers(0, n_rows, size=(n_rows, width)).astype(np.int32)
    idx[:, 0] = np.arange(n_rows)
    ws = [rng.standard_normal((n_rows, width)) for _ in range(n_ops)]
    u = rng.standard_normal((n_rows, F))
    ref = O.spmm(idx, ws, u)
    absum = np.stack([_absum(idx, w, u) for w in ws])
    ell64 = R.EllOperators(ctx, idx, ws, storage=R.F64)
    d_u = ctx.to_device(u)
    d_o = ctx.alloc((n_ops, n_rows, F), np.float64)
    ell64.spmm(F, d_u, d_o)
    got = d_o.download()
    assert rel_apply_err(got, ref, absum) <= TOL64
    ell64.close()
    ell32 = R.EllOperators(ctx, idx, ws, storage=R.F32)
    d_u32 = ctx.to_device(u.astype(np.float32))
    for acc in (R.F64, R.F32):
        d_o32 = ctx.alloc((n_ops, n_rows, F), np.float32)
        ell32.spmm(F, d_u32, d_o32, R.F32, acc)
        dev = rel_apply_err(d_o32.download().astype(np.float64), ref, absum)
        print(f"  fp32 storage, accum {'f64' if acc == R.F64 else 'f32'}: normalised deviation {dev:.2e}")
        assert dev <= TOL32
    ell32.close()


def test_spmm_tuned_shapes_agree(ctx):
    """Every launch shape the autotuner may pick gives bit-identical FP64 results."""
    print("case: spmm launch shapes", flush=True)
    rng = np.random.default_rng(21)
    n_rows, width, F = 3000, 31, 16
    idx = rng.integers(0, n_rows, size=(n_rows, width)).astype(np.int32)
    ws = [rng.standard_normal((n_rows, width)) for _ in range(4)]
    u = rng.standard_normal((n_rows, F))
    ell = R.EllOperators(ctx, idx, ws, storage=R.F64)
    d_u = ctx.to_device(u)
    outs = []
    for fpt in (1, 2, 4, 8):
        for block in (64, 128, 256):
            for batch in ((4, 8, 16) if fpt == 1 else (4, 8)):
                ell.tune_set(0, F, R.F64, R.F64, fpt, block, batch)
                d_o = ctx.alloc((4, n_rows, F), np.float64)
                ell.spmm(F, d_u, d_o)
                outs.append(d_o.download())
                d_o.free()
    R.RBFX.rbfx_tune_clear()
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    ell.close()


@pytest.mark.parametrize("tiles", ["short", "bench"])
@pytest.mark.parametrize("storage", [R.F64, R.F32])
def test_spmm_lds_staged_bit_identical(ctx, c1, storage, tiles):
    """The LDS-staged SpMM shapes (tile records copied into LDS once, 16-bit local indices;
    FP32: also the matrix-core multiply-adds) take the slots in the same order as the direct
    gathers, so the fused 4-operator x 16-field outputs are bit-identical; FP64 also matches the
    CPU restatement to 1e-12 (cancellation-normalised). Tiles: 64 rows capped at 128 nodes
    (many short tiles), or the benchmark's 128-row tiles (48 KB of FP32 records, 78 KB of FP64)."""
    print(f"case: spmm LDS-staged shapes storage={storage} tiles={tiles}", flush=True)
    from rbffd_gpu_amd import tune as TU
    xyz, st = c1["xyz"], c1["st"]
    off, idx = R.csr_from_fixed(st)
    w, _, _ = O.weights(off, idx, xyz, 3.5, R.X | R.Y | R.Z | R.LSFC, dim=3, poly_deg=2)
    ws = [w[t].reshape(st.shape) for t in range(4)]
    F = 16
    rng = np.random.default_rng(4)
    u = rng.standard_normal((xyz.shape[0], F))
    ft = R.F64 if storage == R.F64 else R.F32
    dt = np.float64 if ft == R.F64 else np.float32
    ell = R.EllOperators(ctx, st, ws, storage=storage)
    ell.set_row_order(R.order_morton(xyz))
    if tiles == "short":
        tl, k_max, n_tiles = ell.build_local(64, 128)
    else:
        tl, k_max, n_tiles = ell.build_local(128, (78 if ft == R.F64 else 48) * 1024 // (F * np.dtype(dt).itemsize))
    d_u = ctx.to_device(u.astype(dt))
    outs = {}
    for fpt, block, batch, share in TU.candidates(F, TU.KIND_SPMM, 4, (tl, k_max, n_tiles), 8 if ft == R.F64 else 4,
                                                  ft == R.F32):
        if share not in (4, 5) and block != 256:
            continue
        ell.tune_set(TU.KIND_SPMM, F, ft, ft, fpt, block, batch, share)
        d_o = ctx.alloc((4, xyz.shape[0], F), dt)
        ell.spmm(F, d_u, d_o, ft, ft)
        outs[(fpt, batch, share)] = d_o.download()
        d_o.free()
    R.RBFX.rbfx_tune_clear()
    staged = [k for k in outs if k[2] in (4, 5)]
    assert staged and (ft == R.F64 or any(k[2] == 5 for k in staged))
    base = outs[(4, 8, 0)]
    for k, v in outs.items():
        assert np.array_equal(v, base), k
    if ft == R.F64:
        for t in range(4):
            for f in (0, 7, 15):
                ref = O.apply(off, idx, ws[t].ravel(), u[:, f])
                absum = (np.abs(ws[t]) * np.abs(u[st, f])).sum(1)
                assert rel_apply_err(base[t, :, f], ref, absum) <= 1e-12
    d_u.free()
    ell.close()


def test_ell_rejects_out_of_range_index(ctx):
    rng = np.random.default_rng(1)
    idx = rng.integers(0, 100, size=(100, 5)).astype(np.int32)
    idx[7, 3] = 100
    with pytest.raises(R.RbfxError):
        R.EllOperators(ctx, idx, [np.ones((100, 5))])


def test_f32_storage_rounding_matches_oracle(ctx, c1):
    """FP32 weight storage: the row-sum-preserving rounding is bit-identical to the CPU
    restatement (applied with FP64 fields and accumulation, any differing weight would show
    at the 1e-8 level), RBFX_F32_RN is the plain cast, and the stored row sums of the
    derivative operators stay near zero."""
    print("case: fp32 weight rounding", flush=True)
    xyz, st = c1["xyz"], c1["st"]
    off, idx = R.csr_from_fixed(st)
    w, _, _ = O.weights(off, idx, xyz, 3.5, R.X | R.Y | R.Z | R.LSFC, dim=3, poly_deg=2, diag=True)
    ws = [w[t].reshape(st.shape) for t in range(4)]
    rng = np.random.default_rng(3)
    u = rng.standard_normal(st.shape[0])
    d_u = ctx.to_device(u)
    ones = ctx.to_device(np.ones(st.shape[0]))
    d_o = ctx.alloc(st.shape[0], np.float64)
    for storage, rounder in ((R.F32, O.round_f32_rowsum), (R.F32_RN, lambda a: a.astype(np.float32))):
        ell = R.EllOperators(ctx, st, ws, storage=storage)
        for o in range(4):
            w32 = rounder(ws[o]).astype(np.float64)
            ell.spmv(o, d_u, d_o, R.F64, R.F64)
            ref = O.apply(off, idx, w32.reshape(-1), u)
            assert rel_apply_err(d_o.download(), ref, _absum(st, w32, u)) <= TOL64, (storage, o)
            ell.spmv(o, ones, d_o, R.F64, R.F64)
            rs = np.abs(d_o.download())
            if storage == R.F32:
                near = np.abs(ws[o].astype(np.float32).astype(np.float64).sum(1))
                print(f"  op {o}: max |row sum| stored {rs.max():.2e}, nearest rounding {near.max():.2e}")
                assert rs.max() <= 0.1 * near.max()
        ell.close()


def test_row_order_bit_identical(ctx):
    """A processing order (Morton locality order or a random permutation) changes which rows run
    together, not the results: SpMM, single-operator SpMV and the strided SpMV are bit-identical
    to the node-order layout, and a non-permutation is rejected."""
    print("case: row order", flush=True)
    rng = np.random.default_rng(21)
    xyz = R.nodes_fibonacci(3000, 1e-3, 4)
    st = R.knn_kdtree(xyz, 31)
    n_rows, width = st.shape
    ws = [rng.standard_normal((n_rows, width)) for _ in range(3)]
    F = 8
    u = rng.standard_normal((n_rows, F))
    d_u = ctx.to_device(u)
    d_u1 = ctx.to_device(u[:, 0].copy())
    base = R.EllOperators(ctx, st, ws, storage=R.F64)
    d_ref = ctx.alloc((3, n_rows, F), np.float64)
    base.spmm(F, d_u, d_ref)
    ref = d_ref.download()
    d_ref1 = ctx.alloc(n_rows, np.float64)
    base.spmv(1, d_u1, d_ref1)
    ref1 = d_ref1.download()
    ident = np.arange(n_rows, dtype=np.int32)
    for order in (R.order_morton(xyz), rng.permutation(n_rows).astype(np.int32), ident):
        ell = R.EllOperators(ctx, st, ws, storage=R.F64)
        ell.set_row_order(order)
        # a second re-layout starts from the first; ending on the identity drops the order array
        ell.set_row_order(ident if order is ident else order[::-1].copy())
        d_o = ctx.alloc((3, n_rows, F), np.float64)
        ell.spmm(F, d_u, d_o)
        assert np.array_equal(d_o.download(), ref)
        d_o1 = ctx.alloc(n_rows, np.float64)
        ell.spmv(1, d_u1, d_o1)
        assert np.array_equal(d_o1.download(), ref1)
        ell.close()
    single = R.EllOperators(ctx, st, [ws[2]], storage=R.F64)
    single.set_row_order(R.order_morton(xyz))
    d_o1 = ctx.alloc(n_rows, np.float64)
    single.spmv(0, d_u1, d_o1)
    base.spmv(2, d_u1, d_ref1)
    assert np.array_equal(d_o1.download(), d_ref1.download())
    bad = np.arange(n_rows, dtype=np.int32)
    bad[5] = 4
    with pytest.raises(R.RbfxError):
        single.set_row_order(bad)
    single.close()
    base.close()
