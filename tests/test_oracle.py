"""Pins the CPU restatement (oracle/) with checks that do not retype its formulas:
closed forms, finite differences, library routines (LAPACK via scipy, Laguerre polynomials,
sympy), known answers from the reference's own sources, and invariants."""
import os

import numpy as np
import pytest

import oracle as O
import rbffd_gpu_amd as R
from conftest import GOLDEN

U64 = 2.220446049250313e-16


# ---- Gaussian RBF primitives (rbf_gaussian.h) -------------------------------------------------
def test_gaussian_first_derivative_by_finite_differences():
    eps = 1.7
    rng = np.random.default_rng(0)
    for _ in range(20):
        d = rng.uniform(-0.8, 0.8, 3)
        h = 1e-6
        fd = (O.lib().orc_rbf_eval(eps, d[0] + h, d[1], d[2]) - O.lib().orc_rbf_eval(eps, d[0] - h, d[1], d[2])) / (2 * h)
        assert abs(O.lib().orc_rbf_xderiv(eps, *d) - fd) < 1e-8


def test_gaussian_laplacian_by_finite_differences():
    eps = 1.3
    rng = np.random.default_rng(1)
    f = lambda p: O.lib().orc_rbf_eval(eps, *p)
    h = 1e-4
    for dim in (1, 2, 3):
        for _ in range(10):
            d = rng.uniform(-0.7, 0.7, 3)
            lap = 0.0
            for k in range(dim):
                e = np.zeros(3)
                e[k] = h
                lap += (f(d + e) - 2 * f(d) + f(d - e)) / h ** 2
            assert abs(O.lib().orc_rbf_lapl(eps, dim, *d) - lap) < 2e-6, (dim, d)


def test_gaussian_radial_derivatives():
    """phi'(r)/r and phi''(r) of phi(r) = exp(-(eps r)^2) by finite differences in r."""
    eps = 2.1
    for r in (0.05, 0.3, 0.9):
        d = np.array([r, 0.0, 0.0])
        phi = lambda s: np.exp(-(eps * s) ** 2)
        h = 1e-5
        d1 = (phi(r + h) - phi(r - h)) / (2 * h)
        d2 = (phi(r + h) - 2 * phi(r) + phi(r - h)) / h ** 2
        assert abs(O.lib().orc_rbf_rderiv_over_r(eps, *d) - d1 / r) < 1e-7
        assert abs(O.lib().orc_rbf_rderiv2(eps, *d) - d2) < 1e-5


def test_hyperviscosity_polynomial_is_generalised_laguerre():
    """p_k(r) = (-4)^k k! L_k^{(d/2-1)}(eps^2 r^2), the identity stated at rbf_gaussian.h:193-199."""
    from math import factorial

    from scipy.special import eval_genlaguerre

    for dim in (1, 2, 3):
        for k in range(0, 7):
            for eps, r2 in ((1.0, 0.3), (3.5, 0.01), (0.7, 2.0)):
                want = (-4.0) ** k * factorial(k) * eval_genlaguerre(k, dim / 2.0 - 1.0, eps * eps * r2)
                got = O.lib().orc_hv_pk(eps, dim, k, r2)
                assert abs(got - want) <= 1e-9 * max(1.0, abs(want)), (dim, k, got, want)


def test_hyperviscosity_equals_iterated_laplacian_symbolically():
    """Delta^k exp(-eps^2 r^2) in d dimensions (radial Laplacian f'' + (d-1)/r f') by sympy."""
    import sympy as sp

    r, e = sp.symbols("r e", positive=True)
    for dim in (2, 3):
        f = sp.exp(-(e * r) ** 2)
        for k in range(1, 4):
            f = sp.simplify(sp.diff(f, r, 2) + (dim - 1) / r * sp.diff(f, r))
            for ev, rv in ((1.5, 0.4), (3.0, 0.2)):
                want = float(f.subs({e: ev, r: rv}))
                got = O.lib().orc_rbf_hv(ev, dim, k, rv, 0.0, 0.0)
                assert abs(got - want) <= 1e-10 * max(1.0, abs(want)), (dim, k)


# ---- assembly and solve -----------------------------------------------------------------------
def test_system_symmetry_and_solve_against_lapack_golden():
    """The saddle-point matrix is exactly symmetric; the restated LU agrees with LAPACK
    (scipy.linalg.solve) to the forward bound kappa * u on the golden systems."""
    g = np.load(os.path.join(GOLDEN, "lapack_weights.npz"))
    for q in range(int(g["count"])):
        pts, meta, x = g[f"s{q}_xyz"], g[f"s{q}_meta"], g[f"s{q}_x"]
        eps, mask, dim, deg, hvk, cond = meta[0], int(meta[1]), int(meta[2]), int(meta[3]), int(meta[4]), meta[5]
        n = pts.shape[0]
        M, Rr = O.stencil_system(pts, np.arange(n, dtype=np.int32), eps, mask, dim, deg, hvk)
        assert np.array_equal(M, M.T)
        X, info = O.lu_solve(M, Rr)
        assert info == 0
        rel = np.abs(X - x).max() / np.abs(x).max()
        assert rel <= min(1.0, 64 * cond * U64), (q, rel, cond)
        # backward error at the level of machine precision
        r = Rr - M @ X
        be = np.abs(r).max() / (np.abs(M).sum(1).max() * np.abs(X).max() + np.abs(Rr).max())
        assert be < 64 * M.shape[0] * U64


def test_lu_against_lapack_on_random_systems():
    import scipy.linalg as sl

    rng = np.random.default_rng(2)
    for m in (1, 2, 7, 40, 121):
        A = rng.standard_normal((m, m))
        B = rng.standard_normal((m, 3))
        X, info = O.lu_solve(A, B)
        assert info == 0
        assert np.allclose(X, sl.solve(A, B), rtol=1e-9, atol=1e-12)


def test_lu_reports_exact_singularity():
    A = np.array([[1.0, 2.0], [2.0, 4.0]])
    _, info = O.lu_solve(A, np.ones((2, 1)))
    assert info == 2


def test_polynomial_reproduction_degree2():
    """With degree-2 augmentation the X/Y/Z/LAPL weights differentiate quadratics exactly
    (up to rounding amplified by the conditioning)."""
    xyz = R.nodes_fibonacci(2000, 1e-3, 3)
    st = R.knn_kdtree(xyz, 31)
    off, idx = R.csr_from_fixed(st)
    w, _, ns = O.weights(off, idx, xyz, 3.0, R.X | R.Y | R.Z | R.LAPL, dim=3, poly_deg=2)
    assert ns == 0
    w = w.reshape(4, 2000, 31)
    for i in (0, 777, 1999):
        d = xyz[st[i]] - xyz[i]
        px, py, pz = d.T
        scale = np.abs(w[:, i]).sum(1)
        # constants -> 0
        assert np.all(np.abs(w[:, i].sum(1)) < 1e-6 * scale)
        # X applied to x-x0 -> 1, to y -> 0, to (x-x0)^2 -> 0 (at the centre); LAPL of |d|^2 -> 6
        assert abs(w[0, i] @ px - 1) < 1e-6
        assert abs(w[0, i] @ py) < 1e-6 * scale[0]
        assert abs(w[0, i] @ (px * px)) < 1e-6 * scale[0]
        assert abs(w[1, i] @ py - 1) < 1e-6
        assert abs(w[2, i] @ pz - 1) < 1e-6
        assert abs(w[3, i] @ (px * px + py * py + pz * pz) - 6) < 1e-5 * scale[3] / 100 + 1e-6


def test_five_point_laplacian_flat_limit_kat():
    """tests/unified/compute_weights.cxx:313-315: on an 11x11 grid with eps = 0.05, stencil 12
    gives the 5-point Laplacian (-4,1,1,1,1)/h^2 = (-100,25,25,25,25)."""
    g = np.linspace(-1, 1, 11)
    xyz = np.array([[x, y, 0.0] for y in g for x in g])
    st = R.knn_kdtree(xyz, 5)
    off, idx = R.csr_from_fixed(st)
    w, _, ns = O.weights(off, idx, xyz, 0.05, R.LAPL, dim=2, poly_deg=0)
    ws = w[0].reshape(121, 5)[12]
    order = np.argsort(-np.abs(ws))
    assert abs(ws[order[0]] - (-100.0)) < 1e-2
    assert np.allclose(ws[order[1:]], 25.0, atol=1e-2)
    assert st[12][order[0]] == 12


def test_laplace_beltrami_on_spherical_harmonics():
    """Delta_S Y_lm = -l(l+1) Y_lm on nodes exactly on the unit sphere (icosahedral)."""

    xyz = R.nodes_icosahedral(4)
    n = 31
    st = R.knn_kdtree(xyz, n)
    off, idx = R.csr_from_fixed(st)
    N = xyz.shape[0]
    w, _, _ = O.weights(off, idx, xyz, 0.035 * np.sqrt(N) - 0.1, R.LSFC, dim=2, poly_deg=0)
    for l, m in ((1, 0), (2, 1), (3, -2), (4, 3)):
        Y = R.sph_harm(l, m, xyz)
        d = O.apply(off, idx, w[0], Y)
        err = np.abs(d + l * (l + 1) * Y).max() / (l * (l + 1) * np.abs(Y).max())
        assert err < 2e-3, (l, m, err)


def test_surface_gradient_on_a_linear_function():
    """XSFC/YSFC/ZSFC of f = x on the unit sphere is (1 - x^2, -xy, -xz)."""
    xyz = R.nodes_icosahedral(4)
    st = R.knn_kdtree(xyz, 31)
    off, idx = R.csr_from_fixed(st)
    N = xyz.shape[0]
    w, _, _ = O.weights(off, idx, xyz, 0.035 * np.sqrt(N) - 0.1, R.XSFC | R.YSFC | R.ZSFC, dim=3, poly_deg=0)
    x, y, z = xyz.T
    want = [1 - x * x, -x * y, -x * z]
    for t in range(3):
        assert np.abs(O.apply(off, idx, w[t], x) - want[t]).max() < 1e-3


def test_lambda_theta_on_coordinate_functions():
    """(1/cos theta) d/dlambda and d/dtheta of z and x on the sphere."""
    xyz = R.nodes_icosahedral(4)
    st = R.knn_kdtree(xyz, 31)
    off, idx = R.csr_from_fixed(st)
    N = xyz.shape[0]
    w, _, _ = O.weights(off, idx, xyz, 0.035 * np.sqrt(N) - 0.1, R.LAMBDA | R.THETA, dim=2, poly_deg=0)
    x, y, z = xyz.T
    lam, th = np.arctan2(y, x), np.arctan2(z, np.hypot(x, y))
    assert np.abs(O.apply(off, idx, w[0], z)).max() < 1e-3
    assert np.abs(O.apply(off, idx, w[1], z) - np.cos(th)).max() < 1e-3
    assert np.abs(O.apply(off, idx, w[0], x) + np.sin(lam)).max() < 1e-3
    assert np.abs(O.apply(off, idx, w[1], x) + np.sin(th) * np.cos(lam)).max() < 1e-3


# ---- apply, RK4, cosine bell, norms --------------------------------------------------------------
def test_apply_integer_kat():
    rng = np.random.default_rng(3)
    idx = rng.integers(0, 50, size=(50, 7)).astype(np.int32)
    w = rng.integers(-5, 6, size=(50, 7)).astype(float)
    u = rng.integers(-9, 10, size=50).astype(float)
    off, flat = R.csr_from_fixed(idx)
    assert np.array_equal(O.apply(off, flat, w.reshape(-1), u), (w * u[idx]).sum(1))
    out = O.spmm(idx, [w, 2 * w], np.stack([u, -u], 1))
    assert np.array_equal(out[1, :, 1], -2 * (w * u[idx]).sum(1))


def test_rk4_amplification_factor():
    lam = np.linspace(-2.5, 0.8, 64)
    u0 = np.linspace(1, 2, 64)
    idx = np.arange(64, dtype=np.int32)[:, None]
    dt = 0.2
    u1 = O.rk4_step(idx, [lam[:, None]], None, u0[:, None], dt)[:, 0]
    z = lam * dt
    assert np.allclose(u1, u0 * (1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24), rtol=1e-15, atol=0)


def test_cosine_bell_exact_solution():
    a = 1.0
    c = np.array([[0.0, -1.0, 0.0]])  # lambda = 3pi/2, theta = 0: the bell centre
    assert O.cosine_bell_exact(c, a)[0] == pytest.approx(1.0, abs=1e-15)
    far = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    assert np.all(O.cosine_bell_exact(far, a) == 0.0)
    # half height at r = R/2 along the equator: h = (1 + cos(pi/2))/2 = 0.5
    ang = 3 * np.pi / 2 + (a / 3) / 2
    p = np.array([[np.cos(ang), np.sin(ang), 0.0]])
    assert O.cosine_bell_exact(p, a)[0] == pytest.approx(0.5, abs=1e-12)


def test_cosine_bell_velocity_is_solid_body_rotation():
    """alpha = pi/2: the velocity field is rotation about the x axis with speed u0 = 2 pi a / T."""
    xyz = R.nodes_icosahedral(3)
    cl, ct = O.cosine_bell_coefs(xyz, 1.0, np.pi / 2, 1036800.0)
    x, y, z = xyz.T
    lam, th = np.arctan2(y, x), np.arctan2(z, np.hypot(x, y))
    u0 = 2 * np.pi / 1036800.0
    assert np.allclose(-cl, u0 * np.sin(th) * np.cos(lam), atol=1e-20)
    assert np.allclose(-ct, -u0 * np.sin(lam), atol=1e-20)


def test_norms_against_numpy():
    rng = np.random.default_rng(5)
    a, e = rng.standard_normal(100), rng.standard_normal(100)
    n = O.sol_norms(a, e)
    assert n["l1"] == pytest.approx(np.abs(a - e).sum())
    assert n["l2"] == pytest.approx(np.linalg.norm(a - e))
    assert n["linf"] == pytest.approx(np.abs(a - e).max())
    assert n["l2rel"] == pytest.approx(np.linalg.norm(a - e) / np.linalg.norm(e))


def test_planted_mistake_is_caught():
    """A dropped term in the restated RHS (here: the LSFC phi'' term) must fail the
    spherical-harmonic check above: the check is sensitive to the formula."""

    xyz = R.nodes_icosahedral(4)
    st = R.knn_kdtree(xyz, 31)
    off, idx = R.csr_from_fixed(st)
    N = xyz.shape[0]
    eps = 0.035 * np.sqrt(N) - 0.1
    # LSFC restated without its phi'' term equals 1/4 (4 - 3 r^2) phi'/r: build that "wrong"
    # operator from the XSFC-independent pieces by solving with a modified right-hand side
    M_ok = []
    Y = R.sph_harm(2, 1, xyz)
    good = O.apply(off, idx, O.weights(off, idx, xyz, eps, R.LSFC, dim=2)[0][0], Y)
    wrong = np.empty(N)
    for i in range(N):
        M, Rr = O.stencil_system(xyz, st[i], eps, R.LSFC, 2, 0, 0)
        d = xyz[i] - xyz[st[i]]
        r2 = (d * d).sum(1)
        Rr[:31, 0] = 0.25 * (4 - 3 * r2) * (-2 * eps * eps * np.exp(-eps * eps * r2))
        X, _ = O.lu_solve(M, Rr)
        wrong[i] = X[:31, 0] @ Y[st[i]]
    err_good = np.abs(good + 6 * Y).max() / (6 * np.abs(Y).max())
    err_wrong = np.abs(wrong + 6 * Y).max() / (6 * np.abs(Y).max())
    assert err_good < 2e-3 and err_wrong > 1e-1
    del M_ok


def test_sphere_reduced_basis_reproduces_degree2_harmonics():
    """On nodes exactly on the sphere the full degree-2 basis is singular; the 9-function
    basis is not, and Laplace-Beltrami weights then reproduce Delta_S Y_lm = -l(l+1) Y_lm for
    l <= 2 to rounding level (polynomial reproduction)."""

    xyz = R.nodes_icosahedral(4)
    st = R.knn_kdtree(xyz, 31)
    off, idx = R.csr_from_fixed(st)
    N = xyz.shape[0]
    assert O.lib().orc_num_monomials(2 | R.POLY_SPHERE) == 9
    assert O.lib().orc_num_monomials(3 | R.POLY_SPHERE) == 16
    M, _ = O.stencil_system(xyz, st[100], 4.0, R.LSFC, 2, 2, 0)
    assert np.linalg.cond(M) > 1e15  # full basis: singular on the sphere
    M2, _ = O.stencil_system(xyz, st[100], 4.0, R.LSFC, 2, 2 | R.POLY_SPHERE, 0)
    assert np.linalg.cond(M2) < 1e13
    w, _, ns = O.weights(off, idx, xyz, 0.035 * np.sqrt(N) - 0.1, R.LSFC, dim=2, poly_deg=2 | R.POLY_SPHERE)
    assert ns == 0
    for l, m in ((1, 0), (2, -1), (2, 2)):
        Y = R.sph_harm(l, m, xyz)
        d = O.apply(off, idx, w[0], Y)
        assert np.abs(d + l * (l + 1) * Y).max() < 1e-6 * l * (l + 1), (l, m)


def test_round_f32_rowsum_properties():
    """Pins the FP32 row-sum-preserving rounding by what it must satisfy: every value within
    1.5 FP32 ulp of the FP64 weight, zero stays zero, a row whose nearest-rounded values
    already sum exactly is left as the plain cast, and on RBF-FD derivative rows (largest
    weight first) the stored row sum is orders of magnitude closer to the FP64 one."""
    rng = np.random.default_rng(7)
    W = rng.standard_normal((500, 40)) * np.exp(rng.uniform(-6, 6, (500, 40)))
    W[:, -3:] = 0.0  # padding slots
    w32 = O.round_f32_rowsum(W)
    ulp = np.spacing(np.abs(W.astype(np.float32))).astype(np.float64)
    assert np.all(np.abs(w32.astype(np.float64) - W) <= 1.5 * ulp)
    assert np.all(w32[:, -3:] == 0.0)
    Wi = rng.integers(-100, 100, (50, 20)).astype(np.float64) / 8.0  # exact in FP32: no moves
    assert np.array_equal(O.round_f32_rowsum(Wi), Wi.astype(np.float32))
    xyz = R.nodes_fibonacci(2048, 1e-3, 3)
    st = R.knn_kdtree(xyz, 31)
    off, idx = R.csr_from_fixed(st)
    w, _, _ = O.weights(off, idx, xyz, 3.0, R.X | R.LSFC, dim=3, poly_deg=2)
    for t in range(2):
        Wt = w[t].reshape(st.shape)
        S = Wt.sum(1)
        near = np.abs(Wt.astype(np.float32).astype(np.float64).sum(1) - S)
        kept = np.abs(O.round_f32_rowsum(Wt).astype(np.float64).sum(1) - S)
        assert np.median(kept) < 1e-3 * np.median(near)
        # a planted mistake (moving by two ulps instead of one) breaks the 1.5-ulp bound
        bad = O.round_f32_rowsum(Wt).astype(np.float64)
        bad[:, 1] += 2 * np.spacing(np.abs(Wt[:, 1]).astype(np.float32))
        assert np.any(np.abs(bad - Wt) > 1.5 * np.spacing(np.abs(Wt.astype(np.float32))).astype(np.float64))


# ---- degree-3 augmentation, multi-operator RK4, cosine-bell right-hand side -------------------
def _cubics(xyz):
    """Every monomial x^a y^b z^c with a + b + c <= 3 in absolute coordinates, with its exact
    gradient and Laplacian (computed here from the exponents, not from the oracle's tables)."""
    x, y, z = xyz.T
    out = []
    for a in range(4):
        for b in range(4 - a):
            for c in range(4 - a - b):
                def mono(p, q, r):
                    if p < 0 or q < 0 or r < 0:
                        return np.zeros_like(x)
                    return x ** p * y ** q * z ** r
                val = mono(a, b, c)
                grad = [a * mono(a - 1, b, c), b * mono(a, b - 1, c), c * mono(a, b, c - 1)]
                lap = a * (a - 1) * mono(a - 2, b, c) + b * (b - 1) * mono(a, b - 2, c) + c * (c - 1) * mono(a, b, c - 2)
                out.append(((a, b, c), val, grad, lap))
    return out


def test_polynomial_reproduction_degree3_full_basis():
    """Degree-3 augmentation (20 monomials): X/Y/Z/LAPL weights differentiate every cubic
    exactly (up to rounding times the conditioning), in absolute coordinates. A dropped or
    duplicated cubic monomial in the basis fails this (a missing one is not reproduced, a
    duplicate makes the system singular)."""
    xyz = R.nodes_fibonacci(3000, 1e-3, 4)
    st = R.knn_kdtree(xyz, 50)
    off, idx = R.csr_from_fixed(st)
    assert O.lib().orc_num_monomials(3) == 20
    w, _, ns = O.weights(off, idx, xyz, 4.0, R.X | R.Y | R.Z | R.LAPL, dim=3, poly_deg=3)
    assert ns == 0
    w = w.reshape(4, 3000, 50)
    polys = _cubics(xyz)
    assert len(polys) == 20
    for i in (0, 1234, 2999):
        for (abc, val, grad, lap) in polys:
            want = [grad[0][i], grad[1][i], grad[2][i], lap[i]]
            for t in range(4):
                got = w[t, i] @ val[st[i]]
                scale = np.abs(w[t, i]).sum() * np.abs(val[st[i]]).max() + 1.0
                assert abs(got - want[t]) <= 1e-8 * scale, (i, abc, t, got, want[t])


def test_polynomial_reproduction_degree3_sphere_basis():
    """Sphere-reduced degree-3 basis (16 functions) on nodes exactly on the unit sphere:
    Laplace-Beltrami weights reproduce Delta_S p = -l(l+1) p for the harmonic homogeneous
    polynomials of degree l <= 3 (their restrictions are spherical harmonics), and surface
    gradient weights reproduce grad_S p = grad p - l p x (Euler's relation x.grad p = l p)."""
    xyz = R.nodes_icosahedral(4)
    st = R.knn_kdtree(xyz, 50)
    off, idx = R.csr_from_fixed(st)
    N = xyz.shape[0]
    w, _, ns = O.weights(off, idx, xyz, 4 * (0.044 * np.sqrt(N) - 0.14), R.LSFC | R.XSFC | R.YSFC | R.ZSFC, dim=2,
                         poly_deg=3 | R.POLY_SPHERE)
    assert ns == 0
    x, y, z = xyz.T
    harm = [  # (degree, value, gradient)
        (0, np.ones_like(x), [0 * x, 0 * x, 0 * x]),
        (1, z, [0 * x, 0 * x, 1 + 0 * x]),
        (2, x * y, [y, x, 0 * x]),
        (2, x * x - y * y, [2 * x, -2 * y, 0 * x]),
        (3, x * y * z, [y * z, x * z, x * y]),
        (3, x ** 3 - 3 * x * y * y, [3 * x * x - 3 * y * y, -6 * x * y, 0 * x]),
        (3, 2 * z ** 3 - 3 * x * x * z - 3 * y * y * z, [-6 * x * z, -6 * y * z, 6 * z * z - 3 * x * x - 3 * y * y]),
    ]
    for l, p, g in harm:
        lb = O.apply(off, idx, w[0], p)
        assert np.abs(lb + l * (l + 1) * p).max() < 1e-7 * max(1, l * (l + 1)), (l, np.abs(lb + l * (l + 1) * p).max())
        for c in range(3):
            want = g[c] - l * p * xyz[:, c]
            assert np.abs(O.apply(off, idx, w[1 + c], p) - want).max() < 1e-7, (l, c)
    # degree 2 alone does not reproduce the cubic harmonics: the test separates the degrees
    w2, _, _ = O.weights(off, idx, xyz, 4 * (0.044 * np.sqrt(N) - 0.14), R.LSFC, dim=2, poly_deg=2 | R.POLY_SPHERE)
    p = x * y * z
    assert np.abs(O.apply(off, idx, w2[0], p) + 12 * p).max() > 1e-6


def test_rk4_multi_operator_solid_body_rotation():
    """u_t = y D_x u - x D_y u (two operators with coefficient arrays, two fields) is rotation
    about the z axis: u(t, x, y, z) = u0(x cos t + y sin t, -x sin t + y cos t, z). With
    degree-2 weights the operators reproduce the quadratic fields xz and x^2 - y^2 up to the
    rounding of the weights (about kappa u, 1e-6 here), and RK4 adds its O(dt^4) truncation
    error; halving dt changes the result by less than that bound. Swapping the coefficients
    between the operators, dropping one, or flipping a sign gives O(1) errors."""
    xyz = R.nodes_fibonacci(2000, 1e-3, 6)
    st = R.knn_kdtree(xyz, 31)
    off, idx = R.csr_from_fixed(st)
    w, _, ns = O.weights(off, idx, xyz, 3.0, R.X | R.Y, dim=3, poly_deg=2)
    assert ns == 0
    x, y, z = xyz.T

    def exact(t):
        xr, yr = x * np.cos(t) + y * np.sin(t), -x * np.sin(t) + y * np.cos(t)
        return np.column_stack([xr * z, xr * xr - yr * yr])

    T, steps = 1.0, 100
    u = O.rk4_step(st, [w[0], w[1]], [y, -x], exact(0.0), T / steps, n_steps=steps)
    err = np.abs(u - exact(T)).max()
    assert err < 2e-5, err
    u2 = O.rk4_step(st, [w[0], w[1]], [y, -x], exact(0.0), T / (2 * steps), n_steps=2 * steps)
    assert np.abs(u2 - u).max() < 1e-6  # RK4 truncation at dt = 0.01: about 3e-7
    for ws, cs in (([w[0], w[1]], [-x, y]), ([w[0]], [y]), ([w[0], w[1]], [y, x])):
        with np.errstate(all="ignore"):
            bad = O.rk4_step(st, ws, cs, exact(0.0), T / steps, n_steps=steps)
            assert not np.all(np.abs(bad - exact(T)) <= 1e-2)  # also when the wrong system blows up


def test_cosine_bell_rhs_of_a_smooth_field_is_minus_u_grad_h():
    """CosineBell::solve without hyperviscosity on a smooth field h equals the analytic
    -u.grad h of the solid-body velocity (alpha = pi/2: rotation about the x axis,
    u = Omega x x with Omega = (-u0/a, 0, 0)) to discretisation error."""
    xyz = R.nodes_icosahedral(4)
    st = R.knn_kdtree(xyz, 31)
    off, idx = R.csr_from_fixed(st)
    N = xyz.shape[0]
    eps = 0.035 * np.sqrt(N) - 0.1
    w, _, _ = O.weights(off, idx, xyz, eps, R.LAMBDA | R.THETA | R.HV, dim=2, poly_deg=0, hv_k=4,
                        hv_scale=-0.05 * N ** -4.0)
    x, y, z = xyz.T
    h = np.exp(x) * np.sin(2 * y) + z ** 3
    grad = np.column_stack([np.exp(x) * np.sin(2 * y), 2 * np.exp(x) * np.cos(2 * y), 3 * z * z])
    T = 1036800.0
    u0 = 2 * np.pi / T
    omega = np.array([-u0, 0.0, 0.0])
    vel = np.cross(omega, xyz)
    want = -(vel * grad).sum(1)
    # weight columns in ascending DerType bit order: HV (0x20), LAMBDA (0x40), THETA (0x80)
    got = O.cosine_bell_rhs(off, idx, w[1], w[2], w[0], xyz, h, use_hv=0)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 5e-3, err
    # the velocity's sign or the 1/cos(theta) of the zonal term matter
    wrong = O.cosine_bell_rhs(off, idx, w[1], w[2], w[0], xyz, h, alpha=-np.pi / 2, use_hv=0)
    assert np.abs(wrong - want).max() / np.abs(want).max() > 0.5
