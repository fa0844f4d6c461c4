"""GPU batched weight solves vs the CPU restatement and vs LAPACK-solved golden systems.

Weight parity cannot be elementwise at 1e-12: two partial-pivot LU orderings of a system
with condition number kappa agree only to about kappa * u64 (kappa = 1e8..1e12 here). The
bars are therefore
  * normwise backward error ||R - M W|| / (||M|| ||W|| + ||R||) <= 64 * m * u64, and
  * forward agreement |W_gpu - W_ref| / |W_ref| <= 64 * kappa * u64 (capped at 1),
and the applied derivatives agree with the CPU path to the same forward bound.
"""
import os

import numpy as np
import pytest

import oracle as O
import rbffd_gpu_amd as R
from conftest import GOLDEN

pytestmark = pytest.mark.gpu
U64 = 2.220446049250313e-16


def test_weights_c1_fp64(ctx, c1):
    print("case: weights c1 deg2 X|Y|Z|LSFC", flush=True)
    xyz, st = c1["xyz"], c1["st"]
    off, idx = R.csr_from_fixed(st)
    mask = R.X | R.Y | R.Z | R.LSFC
    w_ref, dg_ref, ns = O.weights(off, idx, xyz, 3.5, mask, dim=3, poly_deg=2, diag=True)
    w, dg = ctx.weights(off, idx, xyz, 3.5, mask, dim=3, poly_deg=2, diag=True)
    m = 31 + 10
    print("  backward error gpu max %.2e  cpu max %.2e" % (dg[:, 0].max(), dg_ref[:, 0].max()))
    assert np.all(dg[:, 3] == R.lib.PATH_F64)
    assert dg[:, 0].max() <= 64 * m * U64
    # forward agreement per stencil against the conditioning of its system
    rel = np.abs(w - w_ref).reshape(4, 4096, 31).max(axis=(0, 2)) / np.abs(w_ref).reshape(4, 4096, 31).max(axis=(0, 2))
    print("  forward rel diff max %.2e median %.2e" % (rel.max(), np.median(rel)))
    conds = np.array([np.linalg.cond(O.stencil_system(xyz, st[i], 3.5, mask, 3, 2, 0)[0]) for i in range(0, 4096, 256)])
    assert np.all(rel[::256] <= np.minimum(1.0, 64 * conds * U64))
    # constant reproduction: sum_j w_j = 0 for every derivative operator
    sums = np.abs(w.reshape(4, 4096, 31).sum(2)) / np.abs(w).reshape(4, 4096, 31).sum(2)
    assert sums.max() < 1e-6


def test_weights_lapack_golden(ctx):
    """Golden systems solved by LAPACK (scipy.linalg.solve): uneven sizes 17..101, degree 0/2/3."""
    g = np.load(os.path.join(GOLDEN, "lapack_weights.npz"))
    for q in range(int(g["count"])):
        pts, meta, x = g[f"s{q}_xyz"], g[f"s{q}_meta"], g[f"s{q}_x"]
        eps, mask, dim, deg, hvk, cond = meta[0], int(meta[1]), int(meta[2]), int(meta[3]), int(meta[4]), meta[5]
        n = pts.shape[0]
        off = np.array([0, n], np.int32)
        idx = np.arange(n, dtype=np.int32)
        w, dg = ctx.weights(off, idx, pts, eps, mask, dim=dim, poly_deg=deg, hv_k=hvk, diag=True)
        ref = x[:n].T
        rel = np.abs(w - ref).max() / np.abs(ref).max()
        bound = min(1.0, 64 * cond * U64)
        assert dg[0, 0] <= 64 * (n + 20) * U64, (q, dg[0])
        assert rel <= bound, (q, rel, cond)


def test_weights_uneven_batch_matches_cpu(ctx):
    """One call over stencils of sizes {17,31,50,75,101} mixed, degree 2 and 3, X/Y/Z/LAPL."""
    print("case: uneven batch", flush=True)
    N = 20000
    xyz = R.nodes_fibonacci(N, 1e-3, 1)
    rng = np.random.default_rng(1)
    sizes = rng.choice([17, 31, 50, 75, 101], size=2000)
    rows = rng.choice(N, size=2000, replace=False)
    knn = {k: R.knn_kdtree(xyz, k) for k in (17, 31, 50, 75, 101)}
    c = {17: (0.026, 0.08), 31: (0.035, 0.1), 50: (0.044, 0.14), 75: (0.051, 0.15), 101: (0.058, 0.16)}
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = np.concatenate([knn[k][r] for k, r in zip(sizes, rows)]).astype(np.int32)
    eps = np.array([c[k][0] * np.sqrt(N) - c[k][1] for k in sizes])
    mask = R.X | R.Y | R.Z | R.LAPL
    for deg in (2, 3):
        w_ref, dg_ref, ns = O.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg, diag=True)
        w, dg = ctx.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg, diag=True)
        print(f"  deg {deg}: gpu backward err max {dg[:, 0].max():.2e}, cpu {dg_ref[:, 0].max():.2e}")
        assert dg[:, 0].max() <= 64 * 121 * U64
        # derivatives of a smooth field through both weight sets agree to the forward bound
        x, y, z = xyz.T
        f = np.sin(x) * np.cos(y) + z
        for t in range(4):
            a = O.apply(off, idx, w[t], f)
            b = O.apply(off, idx, w_ref[t], f)
            scale = np.array([np.abs(w_ref[t, off[s]:off[s + 1]]).sum() for s in range(2000)]) * np.abs(f).max()
            assert np.max(np.abs(a - b) / scale) < 1e-4


def test_weights_f32_ir(ctx, c1):
    """FP32 LU + FP64 refinement: converged stencils meet the FP64 backward-error bar,
    the others are re-solved in FP64 and flagged; the weights match the FP64 path."""
    print("case: weights fp32 + IR", flush=True)
    xyz, st = c1["xyz"], c1["st"]
    off, idx = R.csr_from_fixed(st)
    mask = R.X | R.Y | R.Z | R.LSFC
    w64, _ = ctx.weights(off, idx, xyz, 3.5, mask, dim=3, poly_deg=2, diag=True)
    w32, dg = ctx.weights(off, idx, xyz, 3.5, mask, dim=3, poly_deg=2, diag=True, mode=R.SOLVE_F32_LU_F64_IR)
    paths = dg[:, 3].astype(int)
    n_ir = int((paths == R.lib.PATH_F32_IR).sum())
    n_fb = int((paths == R.lib.PATH_F32_FALLBACK_F64).sum())
    print(f"  converged in FP32+IR: {n_ir}, FP64 fallback: {n_fb}")
    assert n_ir + n_fb == 4096
    assert np.all(np.isfinite(w32))
    assert dg[:, 0].max() <= 64 * 41 * U64
    rel = np.abs(w32 - w64).max() / np.abs(w64).max()
    assert rel < 1e-4


def test_weights_sphere_basis_matches_cpu(ctx):
    """Nodes exactly on the sphere, degree 2 and 3 in the sphere-reduced basis, X/Y/Z/HV."""
    xyz = R.nodes_icosahedral(4)
    N = xyz.shape[0]
    st = R.knn_kdtree(xyz, 50)
    off, idx = R.csr_from_fixed(st)
    mask = R.X | R.Y | R.Z | R.HV
    for deg in (2 | R.POLY_SPHERE, 3 | R.POLY_SPHERE):
        w_ref, dg_ref, ns = O.weights(off, idx, xyz, 4 * (0.044 * np.sqrt(N) - 0.14), mask, dim=2, poly_deg=deg, hv_k=4,
                                      hv_scale=-2.0 ** -6 * N ** -4.0, diag=True)
        w, dg = ctx.weights(off, idx, xyz, 4 * (0.044 * np.sqrt(N) - 0.14), mask, dim=2, poly_deg=deg, hv_k=4,
                            hv_scale=-2.0 ** -6 * N ** -4.0, diag=True)
        assert ns == 0
        assert dg[:, 0].max() <= 64 * 66 * U64
        rel = np.abs(w - w_ref).max(axis=1) / np.abs(w_ref).max(axis=1)
        conds = [np.linalg.cond(O.stencil_system(xyz, st[i], 4 * (0.044 * np.sqrt(N) - 0.14), mask, 2, deg, 4)[0])
                 for i in range(0, N, 97)]
        print(f"  deg {deg}: forward rel diff per type {rel}, max cond {max(conds):.1e}")
        assert rel.max() <= min(1.0, 64 * max(conds) * U64)


def test_weights_singular_reported(ctx):
    """A duplicated node makes the system singular: reported, not silently propagated."""
    pts = np.array([[0, 0, 1], [0.1, 0, 0.99], [0.1, 0, 0.99], [0, 0.1, 0.99], [-0.1, 0, 0.99]], float)
    off = np.array([0, 5], np.int32)
    with pytest.raises(R.RbfxError) as e:
        ctx.weights(off, np.arange(5, dtype=np.int32), pts, 1.0, R.X, dim=3, poly_deg=0)
    assert e.value.code == -5


def test_weights_bad_index_rejected(ctx):
    """A stencil index outside the node range is an RBFX_E_INVALID error naming the index,
    raised before any solve runs; a valid call on the same context works afterwards."""
    print("case: out-of-range stencil index", flush=True)
    pts = R.nodes_fibonacci(64, 1e-3, 0)
    off = np.array([0, 5, 10], np.int32)
    for badv in (64, -1):
        idx = np.arange(10, dtype=np.int32)
        idx[7] = badv
        with pytest.raises(R.RbfxError) as e:
            ctx.weights(off, idx, pts, 1.0, R.X, dim=3, poly_deg=0)
        assert e.value.code == -1 and str(badv) in str(e.value)
    w, _ = ctx.weights(off, np.arange(10, dtype=np.int32), pts, 1.0, R.X, dim=3, poly_deg=0)
    assert np.isfinite(w).all()


def test_weights_register_kernel_bit_identical(ctx, monkeypatch):
    """The three FP64 solvers - the persistent matrix-core blocked LU (the default), the
    per-class register-resident elimination and the one-wave LDS elimination - give
    bit-identical weights and diagnostics (the same pivots and the same per-element operation
    order) over every layout (systems of 27..121 unknowns, 4, 5 and 13 right-hand sides), and
    all report an exactly singular large system."""
    print("case: matrix-core LU vs register vs LDS kernel", flush=True)
    ALL_TYPES = (R.X | R.Y | R.Z | R.LAPL | R.R | R.HV | R.LAMBDA | R.THETA | R.LSFC | R.XSFC | R.YSFC | R.ZSFC |
                 R.INTERP)  # 13 right-hand sides: m + T up to 134
    N = 20000
    xyz = R.nodes_fibonacci(N, 1e-3, 1)
    rng = np.random.default_rng(3)
    ks = (17, 31, 50, 54, 75, 101)
    sizes = rng.choice(ks, size=900)
    rows = rng.choice(N, size=900, replace=False)
    knn = {k: R.knn_kdtree(xyz, k) for k in ks}
    c = {17: (0.026, 0.08), 31: (0.035, 0.1), 50: (0.044, 0.14), 54: (0.044, 0.14), 75: (0.051, 0.15),
         101: (0.058, 0.16)}
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = np.concatenate([knn[k][r] for k, r in zip(sizes, rows)]).astype(np.int32)
    eps = np.array([c[k][0] * np.sqrt(N) - c[k][1] for k in sizes])
    got = {}
    for kern in ("lds", "reg", "lu"):
        if kern == "lu":
            monkeypatch.delenv("RBFX_WEIGHTS_F64", raising=False)
        else:
            monkeypatch.setenv("RBFX_WEIGHTS_F64", kern)
        print("  solver", kern, flush=True)
        for mask, deg in ((R.X | R.Y | R.Z | R.LAPL, 2), (R.X | R.Y | R.Z | R.LAPL | R.HV, 3), (ALL_TYPES, 3)):
            got[kern, mask] = ctx.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg, hv_k=2, hv_scale=-1e-3,
                                          diag=True)
        # 101 distinct nodes in the plane z = 0.5: the monomial z is exactly zero at every
        # node, so M has a zero column and elimination meets an exact zero pivot
        plane = np.column_stack([rng.uniform(-0.2, 0.2, (101, 2)), np.full(101, 0.5)])
        with pytest.raises(R.RbfxError) as e:
            ctx.weights(np.array([0, 101], np.int32), np.arange(101, dtype=np.int32), plane, 5.0, R.X, dim=3,
                        poly_deg=2)
        assert e.value.code == -5
    for kern, mask in got:
        if kern == "lds":
            continue
        w, dg = got[kern, mask]
        assert np.isfinite(w).all()
        assert np.array_equal(w, got["lds", mask][0]), (kern, mask)
        assert np.array_equal(dg, got["lds", mask][1]), (kern, mask)
        if mask != ALL_TYPES:  # the angular operators are singular at the poles' stencils
            assert dg[:, 0].max() <= 64 * 121 * U64


def c3_mix(N, n_st, seed, ks=(17, 31, 50, 75, 101)):
    """A sample of the uneven-stencil workload: n_st stencils of the N-node jittered sphere,
    sizes drawn uniformly from ks, eps from the per-size table."""
    xyz = R.nodes_fibonacci(N, 1e-3, 0)
    rng = np.random.default_rng(seed)
    sizes = rng.choice(ks, size=n_st)
    rows = rng.choice(N, size=n_st, replace=False)
    knn = R.knn_kdtree(xyz, max(ks))
    c = {17: (0.026, 0.08), 31: (0.035, 0.1), 50: (0.044, 0.14), 75: (0.051, 0.15), 101: (0.058, 0.16)}
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = np.concatenate([knn[r, :k] for k, r in zip(sizes, rows)]).astype(np.int32)
    eps = np.array([c[k][0] * np.sqrt(N) - c[k][1] for k in sizes])
    return xyz, off, idx, eps, sizes


def test_weights_timed_path_matches_diagnostic_and_oracle(ctx):
    """The production call (no diagnostics: the path bench.py times) gives the same weights,
    bit for bit, as the call with diagnostics, on the uneven 17..101 mix at degree 2 and 3;
    those weights meet the backward-error bar and agree with the CPU restatement's weights
    within the forward bound 64 kappa u of each system."""
    print("case: timed path vs diagnostic path, uneven mix", flush=True)
    xyz, off, idx, eps, sizes = c3_mix(200000, 3000, 5)
    mask = R.X | R.Y | R.Z | R.LAPL
    for deg in (2, 3):
        w_t, none = ctx.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg)
        assert none is None
        w_d, dg = ctx.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg, diag=True)
        assert np.array_equal(w_t, w_d), deg
        mmax = 101 + (10 if deg == 2 else 20)
        print(f"  deg {deg}: backward error max {dg[:, 0].max():.2e}", flush=True)
        assert dg[:, 0].max() <= 64 * mmax * U64
        w_ref, _, ns = O.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg)
        assert ns == 0
        for s in range(0, 3000, 97):
            a, b = off[s], off[s + 1]
            M = O.stencil_system(xyz, idx[a:b], eps[s], mask, 3, deg, 0)[0]
            kappa = np.linalg.cond(M)
            ref = w_ref[:, a:b]
            rel = np.abs(w_t[:, a:b] - ref).max() / np.abs(ref).max()
            assert rel <= min(1.0, 64 * kappa * U64), (deg, s, rel, kappa)


def test_weights_f32_ir_uneven_mix(ctx):
    """FP32 LU + FP64 refinement on the uneven 17..101 mix (degree 2 and 3): the stencils it
    reports as converged meet the FP64 backward-error bar and agree with the oracle within
    64 kappa u; the stencils that fell back to FP64 are bit-identical to the FP64 solve."""
    print("case: fp32 + IR, uneven mix", flush=True)
    xyz, off, idx, eps, sizes = c3_mix(200000, 3000, 6)
    mask = R.X | R.Y | R.Z | R.LAPL
    for deg in (2, 3):
        w64, _ = ctx.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg)
        w32, dg = ctx.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg, diag=True, mode=R.SOLVE_F32_LU_F64_IR)
        paths = dg[:, 3].astype(int)
        conv = np.nonzero(paths == R.lib.PATH_F32_IR)[0]
        fb = np.nonzero(paths == R.lib.PATH_F32_FALLBACK_F64)[0]
        print(f"  deg {deg}: converged {conv.size}, FP64 fallback {fb.size}", flush=True)
        assert conv.size + fb.size == off.size - 1 and conv.size > 0 and fb.size > 0
        mmax = 101 + (10 if deg == 2 else 20)
        assert dg[conv, 0].max() <= 64 * mmax * U64
        for s in fb:
            assert np.array_equal(w32[:, off[s]:off[s + 1]], w64[:, off[s]:off[s + 1]]), s
        w_ref, _, _ = O.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg)
        for s in conv[::max(1, conv.size // 40)]:
            a, b = off[s], off[s + 1]
            kappa = np.linalg.cond(O.stencil_system(xyz, idx[a:b], eps[s], mask, 3, deg, 0)[0])
            ref = w_ref[:, a:b]
            rel = np.abs(w32[:, a:b] - ref).max() / np.abs(ref).max()
            assert rel <= min(1.0, 64 * kappa * U64), (deg, s, rel, kappa)


def test_weights_uneven_full_size_properties(ctx):
    """The uneven-stencil workload at full size (10^6 stencils, sizes uniform over
    {17, 31, 50, 75, 101}, X/Y/Z/LAPL, degree 2 and 3, the production FP64 solve), checked
    through properties that do not need a CPU solve of every system: every stencil meets the
    backward-error bar, and the X and Laplacian weights reproduce the derivatives of a
    polynomial of the augmentation degree (quadratic; cubic where the stencil has room for the
    cubic monomials) to 1e-9 of the magnitude of the terms summed."""
    print("case: uneven workload at full size, 10^6 stencils", flush=True)
    N = 1000000
    ks = np.array([17, 31, 50, 75, 101])
    xyz = R.nodes_fibonacci(N, 1e-3, 0)
    knn = R.knn_kdtree(xyz, 101)
    sizes = np.random.default_rng(1).choice(ks, size=N)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = knn[np.arange(101)[None, :] < sizes[:, None]].astype(np.int32)
    del knn
    c = {17: (0.026, 0.08), 31: (0.035, 0.1), 50: (0.044, 0.14), 75: (0.051, 0.15), 101: (0.058, 0.16)}
    eps = np.array([c[k][0] * np.sqrt(N) - c[k][1] for k in ks])[np.searchsorted(ks, sizes)]
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    xc, yc, zc = x, y, z  # stencil s is centred on node s (row 0 of its kd-tree list)
    assert np.array_equal(idx[off[:-1]], np.arange(N))
    quad = 1 + 2 * x - y + 0.5 * z + x * x - x * y + 0.3 * z * z
    quad_dx, quad_lap = 2 + 2 * xc - yc, np.full(N, 2.6)
    cub = quad + x ** 3 + 0.5 * x * y * z - y * y * z
    cub_dx, cub_lap = quad_dx + 3 * xc * xc + 0.5 * yc * zc, quad_lap + 6 * xc - 2 * zc
    starts = off[:-1]
    mask = R.X | R.Y | R.Z | R.LAPL  # columns X, Y, Z, LAPL
    for deg in (2, 3):
        w, dg = ctx.weights(off, idx, xyz, eps, mask, dim=3, poly_deg=deg, diag=True)
        mmax = 101 + (10 if deg == 2 else 20)
        assert (dg[:, 3].astype(int) == R.lib.PATH_F64).all()
        print(f"  deg {deg}: backward error max {dg[:, 0].max():.2e}", flush=True)
        assert dg[:, 0].max() <= 64 * mmax * U64
        # degree 3 needs 20 monomials: 17-node stencils keep degree 2
        use_cub = (sizes >= 20) if deg == 3 else np.zeros(N, bool)
        for col, name in ((0, "x"), (3, "lap")):
            wc = w[col]
            for vals, exact_all, sel in ((quad, quad_dx if name == "x" else quad_lap, ~use_cub),
                                         (cub, cub_dx if name == "x" else cub_lap, use_cub)):
                if not sel.any():
                    continue
                pj = vals[idx]
                got = np.add.reduceat(wc * pj, starts)
                scale = np.add.reduceat(np.abs(wc * pj), starts) + np.abs(exact_all)
                err = (np.abs(got - exact_all) / scale)[sel]
                print(f"  deg {deg} {name}: {'cubic' if vals is cub else 'quadratic'} reproduction "
                      f"max {err.max():.2e} over {sel.sum()} stencils", flush=True)
                assert err.max() <= 1e-9, (deg, name, float(err.max()))
        del w, dg
