"""ctypes binding of liboracle.so, the CPU restatement of the reference hot path.

TEST INFRASTRUCTURE: imported only by tests/, bench.py's cpu_baseline leg and
__graft_entry__.smoke(), where it is the checker. The product never imports it.
"""
import ctypes as C
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_KNN = os.path.join(HERE, "_ref", "ref_knn")

_lib = None
_lib_path = LIB


def use_library(path):
    """Loads the restatement from another build of oracle.cpp (e.g. a -march=native copy made
    for the host that runs the CPU baseline); must be called before the first use."""
    global _lib_path
    if _lib is not None and path != _lib_path:
        raise RuntimeError("oracle library already loaded from " + _lib_path)
    _lib_path = path


def build_native(out_dir):
    """Compiles oracle.cpp with -O3 -march=native for this host into out_dir; returns the
    library path, or None if the compiler is not available."""
    out = os.path.join(out_dir, "liboracle.so")
    p = subprocess.run(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={out}"], capture_output=True, text=True)
    return out if p.returncode == 0 and os.path.exists(out) else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_lib_path):
            raise OSError(f"{_lib_path} missing: run make -C oracle")
        L = C.CDLL(_lib_path)
        d, i32, u32, vp = C.c_double, C.c_int32, C.c_uint32, C.c_void_p
        for name in ("orc_rbf_eval", "orc_rbf_xderiv", "orc_rbf_rderiv_over_r", "orc_rbf_rderiv2"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d, d, d, d]
        L.orc_rbf_lapl.restype = d
        L.orc_rbf_lapl.argtypes = [d, C.c_int, d, d, d]
        L.orc_rbf_hv.restype = d
        L.orc_rbf_hv.argtypes = [d, C.c_int, C.c_int, d, d, d]
        L.orc_hv_pk.restype = d
        L.orc_hv_pk.argtypes = [d, C.c_int, C.c_int, d]
        L.orc_num_monomials.restype = C.c_int
        L.orc_num_monomials.argtypes = [C.c_int]
        L.orc_effective_degree.restype = C.c_int
        L.orc_effective_degree.argtypes = [C.c_int, C.c_int]
        L.orc_stencil_system.restype = C.c_int
        L.orc_stencil_system.argtypes = [vp, vp, i32, d, u32, i32, i32, i32, vp, vp]
        L.orc_lu_solve.restype = C.c_int
        L.orc_lu_solve.argtypes = [i32, i32, vp, vp]
        L.orc_weights.restype = C.c_int
        L.orc_weights.argtypes = [i32, vp, vp, vp, vp, u32, i32, i32, i32, d, vp, vp, i32]
        L.orc_apply.restype = None
        L.orc_apply.argtypes = [i32, vp, vp, vp, vp, vp]
        L.orc_spmm.restype = None
        L.orc_spmm.argtypes = [i32, i32, vp, i32, vp, i32, vp, vp, i32]
        L.orc_rk4_step.restype = None
        L.orc_rk4_step.argtypes = [i32, i32, vp, i32, vp, vp, i32, d, vp, i32]
        L.orc_rk4_step_ws.restype = None
        L.orc_rk4_step_ws.argtypes = [i32, i32, vp, i32, vp, vp, i32, d, vp, vp, i32]
        L.orc_cart2sph.restype = None
        L.orc_cart2sph.argtypes = [d, d, d, C.POINTER(d), C.POINTER(d), C.POINTER(d)]
        L.orc_cosine_bell_coefs.restype = None
        L.orc_cosine_bell_coefs.argtypes = [i32, vp, d, d, d, vp, vp]
        L.orc_cosine_bell_exact.restype = d
        L.orc_cosine_bell_exact.argtypes = [d, d, d, d]
        L.orc_cosine_bell_rhs.restype = None
        L.orc_cosine_bell_rhs.argtypes = [i32, vp, vp, vp, vp, vp, vp, d, d, d, i32, vp, vp]
        L.orc_sol_norms.restype = None
        L.orc_sol_norms.argtypes = [i32, vp, vp, vp]
        L.orc_round_f32_rowsum.restype = None
        L.orc_round_f32_rowsum.argtypes = [i32, i32, vp, vp]
        L.orc_stokes_nnz.restype = C.c_int64
        L.orc_stokes_nnz.argtypes = [i32, i32]
        L.orc_stokes_assemble.restype = None
        L.orc_stokes_assemble.argtypes = [i32, i32, vp, vp, d, vp, vp, vp]
        L.orc_csr_matvec.restype = None
        L.orc_csr_matvec.argtypes = [i32, vp, vp, vp, vp, vp]
        L.orc_ilu0.restype = C.c_int64
        L.orc_ilu0.argtypes = [i32, vp, vp, vp, i32, i32, vp, vp, vp]
        L.orc_ilu0_apply.restype = None
        L.orc_ilu0_apply.argtypes = [i32, vp, vp, vp, vp]
        L.orc_gmres.restype = C.c_int
        L.orc_gmres.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, d, i32, i32, vp, C.POINTER(i32), C.POINTER(d)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def ntypes(mask):
    return bin(mask & 0xFFFF).count("1")


def effective_degree(poly_deg, n):
    return lib().orc_effective_degree(poly_deg, n)


def stencil_system(xyz, st, eps, dertypes, dim=3, poly_deg=0, hv_k=0):
    xyz = np.ascontiguousarray(xyz, np.float64)
    st = np.ascontiguousarray(st, np.int32)
    n = st.size
    m = n + lib().orc_num_monomials(poly_deg)
    T = ntypes(dertypes)
    M = np.zeros((m, m))
    Rr = np.zeros((m, T))
    lib().orc_stencil_system(_p(xyz), _p(st), n, eps, dertypes, dim, poly_deg, hv_k, _p(M), _p(Rr))
    return M, Rr


def lu_solve(M, Rr):
    A = np.array(M, np.float64, order="C")
    B = np.array(Rr, np.float64, order="C")
    if B.ndim == 1:
        B = B[:, None].copy()
    info = lib().orc_lu_solve(A.shape[0], B.shape[1], _p(A), _p(B))
    return B, info


def weights(st_off, st_idx, xyz, eps, dertypes, dim=3, poly_deg=0, hv_k=0, hv_scale=1.0, diag=False, n_threads=0):
    st_off = np.ascontiguousarray(st_off, np.int32)
    st_idx = np.ascontiguousarray(st_idx, np.int32)
    xyz = np.ascontiguousarray(xyz, np.float64)
    n_st = st_off.size - 1
    e = np.asarray(eps, np.float64)
    if e.ndim == 0:
        e = np.full(n_st, float(e))
    e = np.ascontiguousarray(e)
    T = ntypes(dertypes)
    w = np.empty((T, st_idx.size))
    dg = np.empty((n_st, 2)) if diag else None
    ns = lib().orc_weights(n_st, _p(st_off), _p(st_idx), _p(xyz), _p(e), dertypes, dim, poly_deg, hv_k, hv_scale,
                           _p(w), _p(dg) if diag else None, n_threads)
    return w, dg, ns


def apply(st_off, st_idx, w, u):
    st_off = np.ascontiguousarray(st_off, np.int32)
    st_idx = np.ascontiguousarray(st_idx, np.int32)
    w = np.ascontiguousarray(w, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    out = np.empty(st_off.size - 1)
    lib().orc_apply(out.size, _p(st_off), _p(st_idx), _p(w), _p(u), _p(out))
    return out


def spmm(idx, ws, u, n_threads=0):
    idx = np.ascontiguousarray(idx, np.int32)
    n_rows, n = idx.shape
    W = np.ascontiguousarray(np.stack([np.asarray(w, np.float64).reshape(n_rows, n) for w in ws]))
    u = np.ascontiguousarray(u, np.float64)
    F = 1 if u.ndim == 1 else u.shape[1]
    out = np.empty((len(ws), n_rows, F))
    lib().orc_spmm(n_rows, n, _p(idx), len(ws), _p(W), F, _p(u), _p(out), n_threads)
    return out


class RK4Stepper:
    """Repeated RK4 steps (orc_rk4_step_ws) with the weights stacked and the workspace allocated
    once, outside the steps."""

    def __init__(self, idx, ws, coefs, F, n_threads=0):
        self.idx = np.ascontiguousarray(idx, np.int32)
        self.n_rows, self.n = self.idx.shape
        self.W = np.ascontiguousarray(np.stack([np.asarray(w, np.float64).reshape(self.n_rows, self.n) for w in ws]))
        self.n_ops = len(ws)
        self.keep = [None if c is None else np.ascontiguousarray(c, np.float64) for c in (coefs or [None] * len(ws))]
        self.arr = (C.c_void_p * len(ws))(*[None if c is None else c.ctypes.data for c in self.keep])
        self.F = F
        self.work = np.empty(5 * self.n_rows * F)
        self.n_threads = n_threads

    def step(self, u, dt, n_steps=1):
        """Advances u (float64, C-contiguous, n_rows x F) in place."""
        assert u.dtype == np.float64 and u.flags.c_contiguous and u.size == self.n_rows * self.F
        for _ in range(n_steps):
            lib().orc_rk4_step_ws(self.n_rows, self.n, _p(self.idx), self.n_ops, _p(self.W), self.arr, self.F, dt,
                                  _p(u), _p(self.work), self.n_threads)


def rk4_step(idx, ws, coefs, u, dt, n_steps=1, n_threads=0, time_budget=None):
    """n_steps RK4 steps; with time_budget (seconds) also stop once the steps taken so far
    have used it, and return (u, steps, seconds in the step loop) instead of u."""
    import time
    u = np.array(u, np.float64, order="C")
    F = 1 if u.ndim == 1 else u.shape[1]
    st = RK4Stepper(idx, ws, coefs, F, n_threads)
    t0 = time.perf_counter()
    done = 0
    while done < n_steps:
        st.step(u, dt)
        done += 1
        if time_budget is not None and time.perf_counter() - t0 >= time_budget:
            break
    if time_budget is None:
        return u
    return u, done, time.perf_counter() - t0


def cart2sph(x, y, z):
    t, p, r = C.c_double(), C.c_double(), C.c_double()
    lib().orc_cart2sph(x, y, z, C.byref(t), C.byref(p), C.byref(r))
    return t.value, p.value, r.value


def cosine_bell_coefs(xyz, a=1.0, alpha=np.pi / 2, t_rev=1036800.0):
    xyz = np.ascontiguousarray(xyz, np.float64)
    cl = np.empty(xyz.shape[0])
    ct = np.empty(xyz.shape[0])
    lib().orc_cosine_bell_coefs(xyz.shape[0], _p(xyz), a, alpha, t_rev, _p(cl), _p(ct))
    return cl, ct


def cosine_bell_exact(xyz, a=1.0):
    return np.array([lib().orc_cosine_bell_exact(p[0], p[1], p[2], a) for p in np.asarray(xyz, np.float64)])


def cosine_bell_rhs(st_off, st_idx, w_lambda, w_theta, w_hv, xyz, u, a=1.0, alpha=np.pi / 2, t_rev=1036800.0, use_hv=1):
    args = [np.ascontiguousarray(x, t) for x, t in ((st_off, np.int32), (st_idx, np.int32), (w_lambda, np.float64),
                                                    (w_theta, np.float64), (w_hv, np.float64), (xyz, np.float64),
                                                    (u, np.float64))]
    f = np.empty(args[0].size - 1)
    lib().orc_cosine_bell_rhs(f.size, _p(args[0]), _p(args[1]), _p(args[2]), _p(args[3]), _p(args[4]), _p(args[5]), a,
                              alpha, t_rev, use_hv, _p(args[6]), _p(f))
    return f


def sol_norms(approx, exact):
    a = np.ascontiguousarray(approx, np.float64)
    e = np.ascontiguousarray(exact, np.float64)
    out = np.empty(6)
    lib().orc_sol_norms(a.size, _p(a), _p(e), _p(out))
    return dict(zip(("l1", "l1rel", "l2", "l2rel", "linf", "linfrel"), out))


def round_f32_rowsum(w):
    """FP32 copy of row-major weights (n_rows, width) that keeps the FP64 row sums."""
    w = np.ascontiguousarray(w, np.float64)
    out = np.empty(w.shape, np.float32)
    lib().orc_round_f32_rowsum(w.shape[0], w.shape[1], _p(w), _p(out))
    return out


def ref_knn(xyz, k):
    """Stencils from the reference's own kd-tree (oracle/_ref/ref_knn, built from its sources)."""
    if not os.path.exists(REF_KNN):
        raise FileNotFoundError(REF_KNN)
    xyz = np.ascontiguousarray(xyz, np.float64)
    with tempfile.TemporaryDirectory() as td:
        fn, fo = os.path.join(td, "n.bin"), os.path.join(td, "o.bin")
        with open(fn, "wb") as f:
            f.write(struct.pack("<q", xyz.shape[0]))
            f.write(xyz.tobytes())
        subprocess.run([REF_KNN, fn, str(k), fo], check=True, stdout=subprocess.DEVNULL)
        d = np.fromfile(fo, dtype=np.int32)
    return d[2:].reshape(d[0], d[1])


def knn_brute(xyz, k):
    """Exact k nearest by full sort of squared distances (ties by index)."""
    xyz = np.asarray(xyz, np.float64)
    out = np.empty((xyz.shape[0], k), np.int32)
    for i in range(xyz.shape[0]):
        d = ((xyz - xyz[i]) ** 2).sum(1)
        out[i] = np.lexsort((np.arange(d.size), d))[:k]
    return out


# ---- steady Stokes (oracle_stokes.cpp) ----------------------------------------------------

class CSR:
    """CSR matrix with int64 row offsets (indptr), int32 columns (ascending) and FP64 values."""

    def __init__(self, indptr, indices, data, n):
        self.indptr = np.ascontiguousarray(indptr, np.int64)
        self.indices = np.ascontiguousarray(indices, np.int32)
        self.data = np.ascontiguousarray(data, np.float64)
        self.n = n

    def scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.data, self.indices, self.indptr), shape=(self.n, self.n))

    def matvec(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n)
        lib().orc_csr_matvec(self.n, _p(self.indptr), _p(self.indices), _p(self.data), _p(x), _p(y))
        return y


def stokes_assemble(idx, w4, eta=1.0):
    """Steady-Stokes matrix (stokes_steady_pde.cpp:61-230) from fixed-width stencils idx (N, n)
    and the LSFC, XSFC, YSFC, ZSFC weights w4 (4, N, n)."""
    idx = np.ascontiguousarray(idx, np.int32)
    N, n = idx.shape
    w4 = np.ascontiguousarray(w4, np.float64).reshape(4, N, n)
    nnz = lib().orc_stokes_nnz(N, n)
    ip = np.empty(4 * N + 5, np.int64)
    ind = np.empty(nnz, np.int32)
    val = np.empty(nnz)
    lib().orc_stokes_assemble(N, n, _p(idx), _p(w4), eta, _p(ip), _p(ind), _p(val))
    return CSR(ip, ind, val, 4 * N + 4)


def ilu0(A, row_start=0, row_end=None):
    """ILU(0) of the block [row_start, row_end)^2 of A, identity rows elsewhere (ilu0.hpp)."""
    row_end = A.n if row_end is None else row_end
    lip = np.empty(A.n + 1, np.int64)
    lind = np.empty(A.indices.size, np.int32)
    lval = np.empty(A.indices.size)
    nnz = lib().orc_ilu0(A.n, _p(A.indptr), _p(A.indices), _p(A.data), row_start, row_end, _p(lip), _p(lind), _p(lval))
    return CSR(lip, lind[:nnz].copy(), lval[:nnz].copy(), A.n)


def ilu0_apply(F, v):
    v = np.array(v, np.float64, copy=True)
    lib().orc_ilu0_apply(F.n, _p(F.indptr), _p(F.indices), _p(F.data), _p(v))
    return v


def gmres(A, b, F=None, tol=1e-8, max_iter=300, krylov=20):
    """ViennaCL-style GMRES(krylov) with left preconditioner F (an ilu0() factor, or None).
    Returns (x, converged, iterations, error estimate ||M^-1 r|| / ||b||)."""
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty(A.n)
    it, err = C.c_int32(0), C.c_double(0.0)
    z = None
    fp = (_p(F.indptr), _p(F.indices), _p(F.data)) if F is not None else (z, z, z)
    rc = lib().orc_gmres(A.n, _p(A.indptr), _p(A.indices), _p(A.data), *fp, _p(b), tol, max_iter, krylov, _p(x),
                         C.byref(it), C.byref(err))
    return x, rc == 0, it.value, err.value
