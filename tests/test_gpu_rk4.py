"""GPU RK4 time stepping vs the CPU restatement of TimeDependentPDE::advanceRungeKutta4.

FP64 path: max |u_gpu - u_cpu| / max |u_cpu| <= 1e-12 after several steps. The stage
arithmetic is kept unfused so only the derivative sums (fused multiply-adds on the GPU)
differ from the CPU order.
"""
import numpy as np
import pytest

import oracle as O
import rbffd_gpu_amd as R

pytestmark = pytest.mark.gpu


def _c1_ops(c1, mask=R.X | R.Y | R.Z | R.HV, dim=2, hv_k=4, deg=0):
    xyz, st = c1["xyz"], c1["st"]
    off, idx = R.csr_from_fixed(st)
    scale = -5e-2 * 4096.0 ** -hv_k
    w, _, ns = O.weights(off, idx, xyz, 3.5, mask, dim=dim, poly_deg=deg, hv_k=hv_k, hv_scale=scale)
    assert ns == 0
    return xyz, st, [w[t].reshape(st.shape) for t in range(w.shape[0])]


def _fields(xyz, F, seed=0):
    rng = np.random.default_rng(seed)
    x, y, z = xyz.T
    basis = np.stack([np.ones_like(x), x, y, z, x * y, y * z, x * z, x * x - y * y, np.sin(2 * x) * z,
                      np.cos(3 * y), np.exp(z)], 1)
    return basis @ rng.standard_normal((basis.shape[1], F))


@pytest.mark.parametrize("F,graph", [(1, True), (4, True), (16, True), (16, False)])
def test_rk4_fp64_matches_cpu(ctx, c1, F, graph):
    print(f"case: rk4 fp64 F={F} graph={graph}", flush=True)
    xyz, st, ws = _c1_ops(c1)
    om = np.array([0.3, -0.2, 1.0])
    coefs = [-(om[1] * xyz[:, 2] - om[2] * xyz[:, 1]), -(om[2] * xyz[:, 0] - om[0] * xyz[:, 2]),
             -(om[0] * xyz[:, 1] - om[1] * xyz[:, 0]), None]
    u0 = _fields(xyz, F)
    dt = 1e-3
    ref = O.rk4_step(st, ws, coefs, u0, dt, n_steps=5)
    ell = R.EllOperators(ctx, st, ws, storage=R.F64)
    rk = R.RK4(ctx, ell, F, R.F64, R.F64, coefs=coefs, use_graph=graph)
    rk.set_state(u0)
    rk.step(dt, 5)
    got = rk.get_state()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"  rel err {err:.2e}")
    assert err <= 1e-12
    rk.close()
    ell.close()


def test_rk4_graph_equals_eager(ctx, c1):
    xyz, st, ws = _c1_ops(c1)
    u0 = _fields(xyz, 8, seed=3)
    outs = []
    for graph in (True, False):
        ell = R.EllOperators(ctx, st, ws, storage=R.F32)
        rk = R.RK4(ctx, ell, 8, R.F32, R.F64, coefs=[xyz[:, 1], -xyz[:, 0], None, None], use_graph=graph)
        rk.set_state(u0)
        for _ in range(3):  # several replays with host work in between
            rk.step(2e-3, 2)
            ctx.sync()
        outs.append(rk.get_state())
        rk.close()
        ell.close()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("storage", [R.F64, R.F32])
def test_rk4_shared_loads_bit_identical(ctx, c1, storage):
    """The shared index/weight loads (lanes of a row split the loads, DPP broadcast) keep the
    summation order, so every launch shape gives bit-identical states; the FP64 path also
    matches the CPU restatement."""
    print(f"case: rk4 shared loads storage={storage}", flush=True)
    xyz, st, ws = _c1_ops(c1)
    F = 16
    u0 = _fields(xyz, F, seed=5)
    om = np.array([0.3, -0.2, 1.0])
    coefs = [-(om[1] * xyz[:, 2] - om[2] * xyz[:, 1]), -(om[2] * xyz[:, 0] - om[0] * xyz[:, 2]),
             -(om[0] * xyz[:, 1] - om[1] * xyz[:, 0]), None]
    ell = R.EllOperators(ctx, st, ws, storage=storage)
    ell.set_row_order(R.order_morton(xyz))
    tl, k_max, n_tiles = ell.build_local(64, 128)  # the node cap shortens tiles: variable tile sizes
    print(f"  LDS tiles of <= {tl} rows, at most {k_max} distinct nodes, {n_tiles} tiles")
    assert k_max <= 128 and n_tiles > (st.shape[0] + 63) // 64
    rk = R.RK4(ctx, ell, F, storage, R.F64, coefs=coefs)
    outs = {}
    from rbffd_gpu_amd import tune as TU
    for fpt, block, batch, share in TU.candidates(F, TU.KIND_RK4, ell.n_ops, (tl, k_max, n_tiles),
                                                  8 if storage == R.F64 else 4):
        if block != 128 and share != 4:
            continue
        ell.tune_set(TU.KIND_RK4, F, storage, R.F64, fpt, block, batch, share)
        rk.set_state(u0)
        rk.step(1e-3, 3)
        outs[(fpt, batch, share)] = rk.get_state()
    assert any(k[2] for k in outs) and any(k[2] == 4 for k in outs)
    base = outs[(4, 8, 0)]
    for k, v in outs.items():
        assert np.array_equal(v, base), k
    if storage == R.F64:
        ref = O.rk4_step(st, ws, coefs, u0, 1e-3, n_steps=3)
        assert np.abs(base - ref).max() <= 1e-12 * np.abs(ref).max()
    rk.close()
    ell.close()


def test_rk4_amplification_factor(ctx):
    """du/dt = lambda u through a diagonal operator: one step multiplies u by
    R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24, z = lambda dt (closed form of classic RK4)."""
    n = 4096
    lam = np.linspace(-3.0, 1.0, n)
    idx = np.arange(n, dtype=np.int32)[:, None]
    u0 = np.linspace(0.5, 2.0, n)[:, None].copy()
    dt = 0.1
    z = lam * dt
    amp = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    ell = R.EllOperators(ctx, idx, [lam[:, None]], storage=R.F64)
    rk = R.RK4(ctx, ell, 1, R.F64, R.F64)
    rk.set_state(u0)
    rk.step(dt, 1)
    got = rk.get_state()[:, 0]
    assert np.max(np.abs(got - amp * u0[:, 0]) / np.abs(amp * u0[:, 0])) < 1e-14
    rk.close()
    ell.close()


@pytest.mark.parametrize("n_ops,rows", [(3, 64), (4, 64), (4, 128)])
def test_rk4_matrix_core_fp32_bit_identical(ctx, c1, n_ops, rows):
    """FP32 weights, fields and accumulation: the LDS-staged stage kernel whose multiply-adds
    run on the matrix cores (4x4x1 FP32 MFMA, one 4x4 block per row) takes the slots in the
    same order with one rounding per multiply-add, so its states are bit-identical to the
    VALU kernel's; with 3 operators the fourth lane of each row feeds operator 0's weights
    into the discarded fourth accumulator row. 128-row tiles run 512-thread blocks."""
    print(f"case: rk4 matrix-core fp32 n_ops={n_ops} tile rows {rows}", flush=True)
    xyz, st, ws = _c1_ops(c1)
    ws = ws[:n_ops]
    F = 16
    u0 = _fields(xyz, F, seed=13)
    coefs = [xyz[:, 1], -xyz[:, 0], 0.1 * xyz[:, 2], None][:n_ops]
    ell = R.EllOperators(ctx, st, ws, storage=R.F32)
    ell.set_row_order(R.order_morton(xyz))
    tl, k_max, n_tiles = ell.build_local(rows, 2 * rows)  # variable tile sizes, partly filled blocks
    rk = R.RK4(ctx, ell, F, R.F32, R.F32, coefs=coefs)
    from rbffd_gpu_amd import tune as TU
    shapes = [c for c in TU.candidates(F, TU.KIND_RK4, n_ops, (tl, k_max, n_tiles), 4, True) if c[3] in (4, 5)]
    assert sorted({c[3] for c in shapes}) == [4, 5]
    outs = {}
    for fpt, block, batch, share in shapes:
        ell.tune_set(TU.KIND_RK4, F, R.F32, R.F32, fpt, block, batch, share)
        rk.set_state(u0)
        rk.step(1e-3, 3)
        outs[(fpt, batch, share)] = rk.get_state()
    base = outs[(4, 8, 4)]
    for k, v in outs.items():
        print(f"  shape {k}: max |diff| {np.abs(v - base).max():.3e}")
    for k, v in outs.items():
        assert np.array_equal(v, base), k
    # the matrix-core shapes exist only for FP32 accumulation
    with pytest.raises(R.RbfxError):
        ell.tune_set(TU.KIND_RK4, F, R.F32, R.F64, 4, tl * 4, 8, 5)
    R.RBFX.rbfx_tune_clear()
    rk.close()
    ell.close()


def test_rk4_default_shape_uses_lds_tables(ctx, c1):
    """Without a tuning entry, an operator set with LDS staging tables steps with the
    LDS-staged shape (matrix cores for FP32 weights, fields and accumulation), and its states
    equal a direct-gather shape's bit for bit."""
    print("case: rk4 default shape with LDS tables", flush=True)
    from rbffd_gpu_amd import tune as TU
    xyz, st, ws = _c1_ops(c1)
    F = 16
    u0 = _fields(xyz, F, seed=21)
    coefs = [xyz[:, 1], -xyz[:, 0], 0.1 * xyz[:, 2], None]
    for storage, ft, acc, want in ((R.F32, R.F32, R.F32, 5), (R.F32, R.F32, R.F64, 4), (R.F64, R.F64, R.F64, 4)):
        R.RBFX.rbfx_tune_clear()
        ell = R.EllOperators(ctx, st, ws, storage=storage)
        assert ell.tune_get(TU.KIND_RK4, F, ft, acc)[3] == 0  # no tables: direct gathers
        ell.set_row_order(R.order_morton(xyz))
        tl, k_max, _ = ell.build_local(128, 256)
        assert ell.tune_get(TU.KIND_RK4, F, ft, acc) == (4, tl * 4, 4, want)
        rk = R.RK4(ctx, ell, F, ft, acc, coefs=coefs)
        rk.set_state(u0)
        rk.step(1e-3, 2)
        staged = rk.get_state()
        ell.tune_set(TU.KIND_RK4, F, ft, acc, 4, 128, 8, 0)
        rk.set_state(u0)
        rk.step(1e-3, 2)
        assert np.array_equal(staged, rk.get_state()), (storage, ft, acc)
        R.RBFX.rbfx_tune_clear()
        rk.close()
        ell.close()


def test_rk4_fp32_storage_deviation(ctx, c1):
    """FP32-stored weights and stage fields with FP64 state: deviation from the FP64 path."""
    xyz, st, ws = _c1_ops(c1)
    coefs = [xyz[:, 1], -xyz[:, 0], 0.1 * xyz[:, 2], None]
    u0 = _fields(xyz, 16, seed=9)
    res = {}
    for storage, acc in ((R.F64, R.F64), (R.F32, R.F64), (R.F32, R.F32)):
        ell = R.EllOperators(ctx, st, ws, storage=storage)
        rk = R.RK4(ctx, ell, 16, storage, acc, coefs=coefs)
        rk.set_state(u0)
        rk.step(1e-3, 10)
        res[(storage, acc)] = rk.get_state()
        rk.close()
        ell.close()
    ref = res[(R.F64, R.F64)]
    for key in ((R.F32, R.F64), (R.F32, R.F32)):
        dev = np.abs(res[key] - ref).max() / np.abs(ref).max()
        print(f"  {key}: relative deviation from FP64 after 10 steps {dev:.2e}")
        assert dev <= 1e-6


def test_cosine_bell_rhs_and_step(ctx, c1):
    """Cosine bell pieces: GPU velocity coefficients and initial bell vs the CPU restatement,
    then one RK4 step of the bell against orc_rk4_step with the same operators."""
    xyz, st = c1["xyz"], c1["st"]
    d_xyz = ctx.to_device(xyz)
    cl, ct = ctx.cosine_bell_coefs(d_xyz)
    cl_ref, ct_ref = O.cosine_bell_coefs(xyz)
    assert np.max(np.abs(cl.download() - cl_ref)) <= 1e-15 * np.abs(cl_ref).max() * 10
    assert np.max(np.abs(ct.download() - ct_ref)) <= 1e-15 * np.abs(ct_ref).max() * 10
    h0 = ctx.cosine_bell_init(d_xyz).download()[:, 0]
    assert np.max(np.abs(h0 - O.cosine_bell_exact(xyz))) < 1e-14
    off, idx = R.csr_from_fixed(st)
    mask = R.LAMBDA | R.THETA | R.HV
    hv_scale = -5e-2 * 4096.0 ** -4
    w, _, ns = O.weights(off, idx, xyz, 0.035 * 64 - 0.1, mask, dim=2, poly_deg=0, hv_k=4, hv_scale=hv_scale)
    # the generic combination sum_o c_o D_o h equals CosineBell::solve's expression exactly
    f_ref = O.cosine_bell_rhs(off, idx, w[0], w[1], w[2], xyz, h0)
    ws = [w[t].reshape(st.shape) for t in range(3)]
    d = O.spmm(st, ws, h0)
    f_gen = cl_ref * d[0, :, 0] + ct_ref * d[1, :, 0] + d[2, :, 0]
    assert np.array_equal(f_gen, f_ref)
    dt = 1036800.0 / 300
    ref = O.rk4_step(st, ws, [cl_ref, ct_ref, None], h0[:, None], dt, n_steps=3)
    ell = R.EllOperators(ctx, st, ws, storage=R.F64)
    rk = R.RK4(ctx, ell, 1, coefs=[cl, ct, None])
    rk.set_state(h0[:, None])
    rk.step(dt, 3)
    got = rk.get_state()
    assert np.abs(got - ref).max() / np.abs(ref).max() <= 1e-12
    rk.close()
    ell.close()


def test_rk4_row_order_bit_identical(ctx, c1):
    """RK4 on an operator set with a Morton processing order equals the node-order run bit for
    bit, including a graph captured before the re-layout (it is recaptured)."""
    xyz, st, ws = _c1_ops(c1)
    F = 16
    u0 = _fields(xyz, F, seed=8)
    om = np.array([0.3, -0.2, 1.0])
    coefs = [-(om[1] * xyz[:, 2] - om[2] * xyz[:, 1]), -(om[2] * xyz[:, 0] - om[0] * xyz[:, 2]),
             -(om[0] * xyz[:, 1] - om[1] * xyz[:, 0]), None]
    outs = []
    for ordered in (False, True):
        ell = R.EllOperators(ctx, st, ws, storage=R.F32)
        rk = R.RK4(ctx, ell, F, R.F32, R.F64, coefs=coefs)
        rk.set_state(u0)
        rk.step(1e-3, 2)
        if ordered:
            ell.set_row_order(R.order_morton(xyz))
        rk.step(1e-3, 2)
        outs.append(rk.get_state())
        rk.close()
        ell.close()
    assert np.array_equal(outs[0], outs[1])


def test_rk4_full_size_matches_oracle(ctx):
    """The RK4 workload at full size (N = 10^6 Morton-numbered sphere nodes, n = 50 kd-tree
    stencils, surface d/dx, d/dy, d/dz and Delta^4 hyperviscosity x 16 fields, the production
    LDS tiles): two FP64 steps agree with the oracle's FP64 steps on the same weights to 1e-12;
    the FP32 matrix-core stage equals the FP32 VALU stage bit for bit, and deviates from FP64 by
    less than 1e-6 after the two steps."""
    print("case: rk4 at full size, N = 10^6, n = 50, 4 operators x 16 fields", flush=True)
    from rbffd_gpu_amd import tune as TU
    N, n, F = 1000000, 50, 16
    xyz = R.nodes_fibonacci(N, 1e-3, 0)
    xyz /= np.linalg.norm(xyz, axis=1)[:, None]
    xyz = np.ascontiguousarray(xyz[R.order_morton(xyz)])
    st = R.knn_kdtree(xyz, n)
    mask = R.XSFC | R.YSFC | R.ZSFC | R.HV  # columns HV, XSFC, YSFC, ZSFC (ascending bits)
    off, idx = R.csr_from_fixed(st)
    w, _ = ctx.weights(off, idx, xyz, 4.0 * (0.044 * np.sqrt(N) - 0.14), mask, dim=2, poly_deg=2 | R.POLY_SPHERE,
                       hv_k=4, hv_scale=-2.0 ** 9 * float(N) ** -4)
    ws = [w[t].reshape(N, n) for t in range(4)]
    del w
    axis = np.array([0.3, -0.2, 1.0])
    vel = np.cross(2 * np.pi * axis / np.linalg.norm(axis), xyz)
    coefs = [None, -vel[:, 0], -vel[:, 1], -vel[:, 2]]
    u0 = _fields(xyz, F, seed=5)
    dt = 2.5e-4
    ref = O.rk4_step(st, ws, coefs, u0, dt, n_steps=2)
    order = R.order_morton(xyz)
    ell64 = R.EllOperators(ctx, st, ws, storage=R.F64)
    ell64.set_row_order(order)
    ell64.build_local(128, 78 * 1024 // (F * 8))  # the tiles the benchmark times
    rk64 = R.RK4(ctx, ell64, F, R.F64, R.F64, coefs=coefs)
    rk64.set_state(u0)
    rk64.step(dt, 2)
    got64 = rk64.get_state()
    err = np.abs(got64 - ref).max() / np.abs(ref).max()
    print(f"  FP64 vs oracle: {err:.2e}", flush=True)
    assert err <= 1e-12
    R.RBFX.rbfx_tune_clear()
    ell32 = R.EllOperators(ctx, st, ws, storage=R.F32)
    ell32.set_row_order(order)
    tl, k_max, n_tiles = ell32.build_local(128, 48 * 1024 // (F * 4))
    assert ell32.tune_get(TU.KIND_RK4, F, R.F32, R.F32) == (4, tl * 4, 4, 5)  # the matrix-core shape
    rk32 = R.RK4(ctx, ell32, F, R.F32, R.F32, coefs=coefs)
    rk32.set_state(u0)
    rk32.step(dt, 2)
    mx = rk32.get_state()
    ell32.tune_set(TU.KIND_RK4, F, R.F32, R.F32, 4, tl * 4, 4, 4)  # the VALU LDS-staged shape
    rk32.set_state(u0)
    rk32.step(dt, 2)
    assert np.array_equal(mx, rk32.get_state())
    dev = np.abs(mx - got64).max() / np.abs(got64).max()
    print(f"  FP32 (matrix cores) vs FP64 after 2 steps: {dev:.2e}; {n_tiles} tiles of <= {tl} rows, "
          f"<= {k_max} nodes", flush=True)
    assert dev <= 1e-6
    R.RBFX.rbfx_tune_clear()
    for o in (rk32, ell32, rk64, ell64):
        o.close()
