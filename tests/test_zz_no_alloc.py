"""rbfx_weights_noalloc (scratch in a static device array, no device allocation) against
rbfx_weights on the same inputs: bit-identical weights and diagnostics, on two calls in a row
with the second smaller than the first (the shared scratch keeps the first call's contents),
in FP64 and in FP32 + refinement mode; sizes above the scratch maximum are rejected."""
import ctypes as C

import numpy as np
import pytest

import rbffd_gpu_amd as R

pytestmark = pytest.mark.gpu
C3 = {17: (0.026, 0.08), 31: (0.035, 0.1), 50: (0.044, 0.14), 75: (0.051, 0.15), 101: (0.058, 0.16)}


def mix(xyz, knn, n_st, seed, ks):
    N = xyz.shape[0]
    rng = np.random.default_rng(seed)
    sizes = rng.choice(ks, size=n_st)
    rows = rng.choice(N, size=n_st, replace=False)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = np.concatenate([knn[r, :k] for k, r in zip(sizes, rows)]).astype(np.int32)
    eps = np.array([C3[k][0] * np.sqrt(N) - C3[k][1] for k in sizes])
    return off, idx, eps


def test_noalloc_matches_weights_two_calls(ctx):
    N = 100000
    xyz = R.nodes_fibonacci(N, 1e-3, 0)
    knn = R.knn_kdtree(xyz, 101)
    d_xyz = ctx.to_device(xyz)
    mask = R.X | R.Y | R.Z | R.LAPL
    cases = [(20000, 11, (17, 31, 50, 75, 101), 2), (3000, 12, (31, 75, 101), 3)]  # second call smaller
    for n_st, seed, ks, deg in cases:
        off, idx, eps = mix(xyz, knn, n_st, seed, ks)
        print(f"case: noalloc vs weights, {n_st} stencils, degree {deg}", flush=True)
        d_off, d_idx = ctx.to_device(off), ctx.to_device(idx)
        for mode in (R.SOLVE_F64_LU, R.SOLVE_F32_LU_F64_IR):
            w_ref, dg_ref = ctx.weights(d_off, d_idx, d_xyz, eps, mask, dim=3, poly_deg=deg, mode=mode, diag=True)
            w_na, dg_na = ctx.weights(d_off, d_idx, d_xyz, eps, mask, dim=3, poly_deg=deg, mode=mode, diag=True,
                                      noalloc=True)
            assert np.array_equal(w_na, w_ref), (n_st, mode)
            assert np.array_equal(dg_na, dg_ref), (n_st, mode)
            w_na2, _ = ctx.weights(d_off, d_idx, d_xyz, eps, mask, dim=3, poly_deg=deg, mode=mode, noalloc=True)
            assert np.array_equal(w_na2, w_ref), (n_st, mode)
        d_off.free()
        d_idx.free()


def test_noalloc_rejects_sizes_above_the_maximum(ctx):
    n_st = R.lib.WEIGHTS_NOALLOC_MAX_STENCILS + 1
    d_off = ctx.alloc(n_st + 1, np.int32)
    d_idx = ctx.alloc(16, np.int32)
    d_xyz = ctx.alloc((4, 3), np.float64)
    d_eps = ctx.alloc(n_st, np.float64)
    d_w = ctx.alloc(16, np.float64)
    ns = C.c_int32()
    rc = R.lib.RBFX.rbfx_weights_noalloc(ctx.h, n_st, d_off.ptr, d_idx.ptr, 4, d_xyz.ptr, d_eps.ptr, R.X, 3, 0, 0, 1.0,
                                         R.SOLVE_F64_LU, d_w.ptr, None, C.byref(ns))
    assert rc == -7
    assert b"maximum" in R.lib.RBFX.rbfx_last_error(ctx.h)
    for a in (d_off, d_idx, d_xyz, d_eps, d_w):
        a.free()
