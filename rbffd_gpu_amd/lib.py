"""ctypes binding of librbfx.so (the C ABI in include/rbfx.h).

Thin wrappers only: every computation runs in the library (HIP kernels on the GPU, or the
host kd-tree / node generator). Errors from the library raise RbfxError with its message.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librbfx.so")

RBFX_OK = 0
X, Y, Z, LAPL, R, HV, LAMBDA, THETA = 0x1, 0x2, 0x4, 0x8, 0x10, 0x20, 0x40, 0x80
LSFC, XSFC, YSFC, ZSFC, INTERP = 0x100, 0x200, 0x400, 0x800, 0x8000
F64, F32 = 0, 1
F32_RN = 2  # weight storage: FP32 with plain nearest rounding (F32 keeps row sums)
SOLVE_F64_LU, SOLVE_F32_LU_F64_IR = 0, 1
POLY_SPHERE = 16  # poly_deg flag: sphere-reduced monomial basis
WEIGHTS_NOALLOC_MAX_STENCILS = 262144  # scratch capacity of rbfx_weights_noalloc (include/rbfx.h)
PATH_F64, PATH_F32_IR, PATH_F32_FALLBACK_F64, PATH_SINGULAR = 0, 1, 2, 3
ERR_NAMES = {-1: "INVALID", -2: "HIP", -3: "NOMEM", -4: "NODEVICE", -5: "SINGULAR", -6: "STATE", -7: "TOO_LARGE"}


class RbfxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rbfx error {code} ({ERR_NAMES.get(code, '?')}): {msg}")
        self.code = code


class _RK4Desc(C.Structure):
    _fields_ = [
        ("ell", C.c_void_p),
        ("n_fields", C.c_int32),
        ("field_storage", C.c_int32),
        ("accum", C.c_int32),
        ("d_coef", C.c_void_p * 16),
        ("use_graph", C.c_int32),
    ]


_lib = None


def load_library(path=None):
    """Loads librbfx.so once; raises OSError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"librbfx.so not found at {p}: run __graft_entry__.build() (make -C rbffd_gpu_amd)")
    lib = C.CDLL(p)
    vp, i32, u32, u64, f64 = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64, C.c_double
    sig = {
        "rbfx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "rbfx_destroy": (None, [vp]),
        "rbfx_last_error": (C.c_char_p, [vp]),
        "rbfx_sync": (C.c_int, [vp]),
        "rbfx_device_info": (C.c_int, [vp, C.c_char_p, i32, C.POINTER(i32), C.POINTER(i32)]),
        "rbfx_dev_alloc": (C.c_int, [vp, u64, C.POINTER(vp)]),
        "rbfx_dev_free": (C.c_int, [vp, vp]),
        "rbfx_h2d": (C.c_int, [vp, vp, vp, u64]),
        "rbfx_d2h": (C.c_int, [vp, vp, vp, u64]),
        "rbfx_d2d": (C.c_int, [vp, vp, vp, u64]),
        "rbfx_stream_copy": (C.c_int, [vp, vp, vp, u64]),
        "rbfx_timer_start": (C.c_int, [vp]),
        "rbfx_timer_stop": (C.c_int, [vp, C.POINTER(f64)]),
        "rbfx_knn_kdtree": (C.c_int, [i32, vp, i32, vp, i32]),
        "rbfx_knn_gpu": (C.c_int, [vp, i32, vp, i32, vp, vp]),
        "rbfx_knn_gpu_device": (C.c_int, [vp, i32, vp, i32, vp, vp]),
        "rbfx_nodes_fibonacci": (C.c_int, [i32, f64, u64, vp]),
        "rbfx_nodes_icosahedral": (C.c_int, [i32, vp, C.POINTER(i32)]),
        "rbfx_order_morton": (C.c_int, [i32, vp, vp]),
        "rbfx_sph_harm": (C.c_int, [i32, i32, i32, vp, vp]),
        "rbfx_ell_set_row_order": (C.c_int, [vp, vp]),
        "rbfx_ell_build_local": (C.c_int, [vp, i32, i32]),
        "rbfx_ell_local_info": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "rbfx_weights": (C.c_int, [vp, i32, vp, vp, i32, vp, vp, u32, i32, i32, i32, f64, i32, vp, vp, C.POINTER(i32)]),
        "rbfx_weights_noalloc": (C.c_int, [vp, i32, vp, vp, i32, vp, vp, u32, i32, i32, i32, f64, i32, vp, vp,
                                           C.POINTER(i32)]),
        "rbfx_ell_create": (C.c_int, [vp, i32, i32, i32, vp, i32, vp, vp, i32, C.POINTER(vp)]),
        "rbfx_ell_destroy": (None, [vp]),
        "rbfx_ell_info": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "rbfx_spmv": (C.c_int, [vp, i32, vp, vp, i32, i32]),
        "rbfx_spmm": (C.c_int, [vp, i32, vp, vp, i32, i32]),
        "rbfx_rk4_create": (C.c_int, [vp, C.POINTER(_RK4Desc), C.POINTER(vp)]),
        "rbfx_rk4_destroy": (None, [vp]),
        "rbfx_rk4_set_state": (C.c_int, [vp, vp]),
        "rbfx_rk4_get_state": (C.c_int, [vp, vp]),
        "rbfx_rk4_step": (C.c_int, [vp, f64, i32]),
        "rbfx_rk4_stage": (C.c_int, [vp, i32, f64]),
        "rbfx_tune_set": (C.c_int, [vp, i32, i32, i32, i32, i32, i32, i32, i32]),
        "rbfx_tune_get": (C.c_int, [vp, i32, i32, i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                    C.POINTER(i32)]),
        "rbfx_tune_clear": (C.c_int, []),
        "rbfx_cosine_bell_coefs": (C.c_int, [vp, i32, vp, f64, f64, f64, vp, vp]),
        "rbfx_cosine_bell_init": (C.c_int, [vp, i32, vp, f64, vp, i32]),
        "rbfx_stokes_create": (C.c_int, [vp, vp, vp, f64, C.POINTER(vp)]),
        "rbfx_stokes_destroy": (None, [vp]),
        "rbfx_stokes_info": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "rbfx_stokes_apply": (C.c_int, [vp, vp, vp]),
        "rbfx_stokes_factor": (C.c_int, [vp, vp]),
        "rbfx_stokes_precond": (C.c_int, [vp, vp]),
        "rbfx_stokes_gmres": (C.c_int, [vp, vp, vp, f64, i32, i32, i32, C.POINTER(i32), C.POINTER(f64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class _Api:
    def __getattr__(self, name):
        return getattr(load_library(), name)


RBFX = _Api()


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def dertype_bits(mask):
    """DerType bits of mask in ascending order (the weight column order)."""
    return [1 << b for b in range(16) if mask & (1 << b)]


def knn_kdtree(xyz, k, n_threads=0):
    """kd-tree stencils (host C++ in librbfx): int32 array (N, k), row i starts with i."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    out = np.empty((n, k), dtype=np.int32)
    rc = RBFX.rbfx_knn_kdtree(n, _ptr(xyz), k, _ptr(out), n_threads)
    if rc:
        raise RbfxError(rc, "knn_kdtree failed")
    return out


def knn_gpu(ctx, xyz, k):
    """The kd-tree stencils of knn_kdtree computed by the GPU search (bit-identical; rows with
    distance ties are recomputed by the host kd-tree). Returns (int32 (N, k), host rows)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    out = np.empty((n, k), dtype=np.int32)
    n_host = C.c_int32()
    ctx.check(RBFX.rbfx_knn_gpu(ctx.h, n, _ptr(xyz), k, _ptr(out), C.byref(n_host)), "knn_gpu")
    return out, n_host.value


def knn_gpu_device(ctx, xyz, k):
    """knn_gpu with the lists left in device memory: returns (DeviceArray int32 (N, k), host rows)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    d_out = ctx.alloc((n, k), np.int32)
    n_host = C.c_int32()
    ctx.check(RBFX.rbfx_knn_gpu_device(ctx.h, n, _ptr(xyz), k, d_out.ptr, C.byref(n_host)), "knn_gpu_device")
    return d_out, n_host.value


def nodes_fibonacci(n, jitter=1e-3, seed=0):
    out = np.empty((n, 3), dtype=np.float64)
    rc = RBFX.rbfx_nodes_fibonacci(n, jitter, seed, _ptr(out))
    if rc:
        raise RbfxError(rc, "nodes_fibonacci failed")
    return out


def nodes_icosahedral(level):
    n = C.c_int32()
    rc = RBFX.rbfx_nodes_icosahedral(level, None, C.byref(n))
    if rc:
        raise RbfxError(rc, "nodes_icosahedral failed")
    out = np.empty((n.value, 3), dtype=np.float64)
    rc = RBFX.rbfx_nodes_icosahedral(level, _ptr(out), C.byref(n))
    if rc:
        raise RbfxError(rc, "nodes_icosahedral failed")
    return out


def order_morton(xyz):
    """Node order along a 3-D Morton curve (int32 permutation), for EllOperators.set_row_order."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    out = np.empty(xyz.shape[0], np.int32)
    rc = RBFX.rbfx_order_morton(xyz.shape[0], _ptr(xyz), _ptr(out))
    if rc:
        raise RbfxError(rc, "order_morton failed")
    return out


def sph_harm(l, m, xyz):
    """Real spherical harmonic Y_lm at the directions of xyz (N x 3), computed in librbfx."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    out = np.empty(xyz.shape[0])
    rc = RBFX.rbfx_sph_harm(int(l), int(m), xyz.shape[0], _ptr(xyz), _ptr(out))
    if rc:
        raise RbfxError(rc, f"sph_harm({l}, {m}) failed")
    return out


def make_fields(xyz, F, seed=0, lmax=20, terms=6):
    """F seeded fields, each a combination of `terms` real Y_lm with l <= lmax and standard
    normal coefficients; returns (N x F array, [(l, m, c), ...] per field)."""
    rng = np.random.default_rng(seed)
    spec = []
    u = np.zeros((np.asarray(xyz).shape[0], F))
    for f in range(F):
        ls = rng.integers(0, lmax + 1, size=terms)
        ms = [int(rng.integers(-l, l + 1)) for l in ls]
        cs = rng.standard_normal(terms)
        spec.append(list(zip(ls.tolist(), ms, cs.tolist())))
        for l, m, c in spec[-1]:
            u[:, f] += c * sph_harm(int(l), int(m), xyz)
    return u, spec


def csr_from_fixed(st):
    st = np.ascontiguousarray(st, dtype=np.int32)
    n, k = st.shape
    return np.arange(0, (n + 1) * k, k, dtype=np.int32), st.reshape(-1)


class Context:
    def __init__(self, device=0):
        lib = load_library()
        h = C.c_void_p()
        rc = lib.rbfx_create(device, C.byref(h))
        if rc:
            raise RbfxError(rc, lib.rbfx_last_error(None).decode())
        self.h = h
        self._buffers = []

    def check(self, rc, what=""):
        if rc:
            raise RbfxError(rc, (what + ": " if what else "") + load_library().rbfx_last_error(self.h).decode())

    def info(self):
        name = C.create_string_buffer(256)
        ncu, clk = C.c_int32(), C.c_int32()
        self.check(RBFX.rbfx_device_info(self.h, name, 256, C.byref(ncu), C.byref(clk)))
        return {"name": name.value.decode(), "n_cu": ncu.value, "clock_khz": clk.value}

    def sync(self):
        self.check(RBFX.rbfx_sync(self.h), "sync")

    def alloc(self, shape, dtype):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d

    def timer_start(self):
        self.check(RBFX.rbfx_timer_start(self.h))

    def timer_stop(self):
        ms = C.c_double()
        self.check(RBFX.rbfx_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def weights(self, st_off, st_idx, xyz, eps, dertypes, dim=3, poly_deg=0, hv_k=0, hv_scale=1.0,
                mode=SOLVE_F64_LU, diag=False, device_out=False, noalloc=False, d_w=None):
        """Batched weights. Host or device inputs; returns (w[T, nnz], diag[n_st, 4] or None).
        noalloc selects rbfx_weights_noalloc (library scratch in a static device array); d_w
        (optional) is a caller-provided device output of shape (T, nnz)."""
        d_off = st_off if isinstance(st_off, DeviceArray) else self.to_device(np.asarray(st_off, np.int32))
        d_idx = st_idx if isinstance(st_idx, DeviceArray) else self.to_device(np.asarray(st_idx, np.int32))
        d_xyz = xyz if isinstance(xyz, DeviceArray) else self.to_device(np.asarray(xyz, np.float64))
        n_st = d_off.shape[0] - 1
        e = np.asarray(eps, np.float64)
        if e.ndim == 0:
            e = np.full(n_st, float(e))
        d_eps = self.to_device(e)
        T = len(dertype_bits(dertypes))
        nnz = d_idx.size
        if d_w is None:
            d_w = self.alloc((T, nnz), np.float64)
        d_diag = self.alloc((n_st, 4), np.float64) if diag else None
        nsing = C.c_int32()
        fn = RBFX.rbfx_weights_noalloc if noalloc else RBFX.rbfx_weights
        rc = fn(self.h, n_st, d_off.ptr, d_idx.ptr, d_xyz.shape[0], d_xyz.ptr, d_eps.ptr, dertypes, dim, poly_deg, hv_k,
                hv_scale, mode, d_w.ptr, d_diag.ptr if d_diag else None, C.byref(nsing))
        self.check(rc, "weights")
        if device_out:
            return d_w, d_diag
        return d_w.download(), (d_diag.download() if d_diag else None)

    def cosine_bell_coefs(self, d_xyz, a=1.0, alpha=3.141592653589793 / 2, t_rev=1036800.0):
        n = d_xyz.shape[0]
        cl, ct = self.alloc(n, np.float64), self.alloc(n, np.float64)
        self.check(RBFX.rbfx_cosine_bell_coefs(self.h, n, d_xyz.ptr, a, alpha, t_rev, cl.ptr, ct.ptr), "cosine_bell_coefs")
        return cl, ct

    def cosine_bell_init(self, d_xyz, a=1.0, n_fields=1):
        n = d_xyz.shape[0]
        h = self.alloc((n, n_fields), np.float64)
        for f in range(n_fields):
            self.check(RBFX.rbfx_cosine_bell_init(self.h, n, d_xyz.ptr, a, C.c_void_p(h.ptr.value + 8 * f), n_fields),
                       "cosine_bell_init")
        return h

    def close(self):
        for b in self._buffers:
            b.free()
        self._buffers = []
        if self.h:
            RBFX.rbfx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceArray:
    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(RBFX.rbfx_dev_alloc(ctx.h, self.nbytes, C.byref(p)), "alloc")
        self._p = p
        ctx._buffers.append(self)

    @property
    def ptr(self):
        return self._p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        self.ctx.check(RBFX.rbfx_h2d(self.ctx.h, self._p, _ptr(a), self.nbytes), "h2d")

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        self.ctx.check(RBFX.rbfx_d2h(self.ctx.h, _ptr(out), self._p, self.nbytes), "d2h")
        return out

    def free(self):
        if self._p:
            RBFX.rbfx_dev_free(self.ctx.h, self._p)
            self._p = C.c_void_p()

    def view(self, offset_elems, shape):
        """Non-owning view of part of this buffer (e.g. one operator's weights)."""
        return DeviceView(self, offset_elems, shape)


class DeviceView:
    def __init__(self, base, offset_elems, shape):
        self.ctx = base.ctx
        self.base = base
        self.dtype = base.dtype
        self.shape = tuple(shape)
        self.size = int(np.prod(self.shape))
        self.nbytes = self.size * self.dtype.itemsize
        assert offset_elems + self.size <= base.size
        self._p = C.c_void_p(base.ptr.value + offset_elems * self.dtype.itemsize)

    @property
    def ptr(self):
        return self._p

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        self.ctx.check(RBFX.rbfx_d2h(self.ctx.h, _ptr(out), self._p, self.nbytes), "d2h")
        return out


class EllOperators:
    """Fused ELL operator set: n_ops operators sharing one (n_rows, width) index array."""

    def __init__(self, ctx, idx, weights, n_cols=None, storage=F64, types=None):
        self.ctx = ctx
        d_idx = idx if isinstance(idx, (DeviceArray, DeviceView)) else ctx.to_device(np.asarray(idx, np.int32))
        n_rows, width = d_idx.shape
        d_ws = [w if isinstance(w, (DeviceArray, DeviceView)) else ctx.to_device(np.asarray(w, np.float64).reshape(n_rows, width))
                for w in weights]
        n_ops = len(d_ws)
        ptrs = (C.c_void_p * n_ops)(*[w.ptr.value for w in d_ws])
        tarr = np.array(types if types is not None else [0] * n_ops, dtype=np.uint32)
        h = C.c_void_p()
        rc = RBFX.rbfx_ell_create(ctx.h, n_rows, n_cols or n_rows, width, d_idx.ptr, n_ops, _ptr(tarr), ptrs, storage,
                                  C.byref(h))
        ctx.check(rc, "ell_create")
        self.h = h
        self.n_rows, self.width, self.n_ops = n_rows, width, n_ops
        self.storage = F32 if storage == F32_RN else storage

    def spmv(self, op, d_u, d_out, field_type=F64, accum=F64):
        self.ctx.check(RBFX.rbfx_spmv(self.h, op, d_u.ptr, d_out.ptr, field_type, accum), "spmv")

    def spmm(self, n_fields, d_u, d_out, field_type=F64, accum=F64):
        self.ctx.check(RBFX.rbfx_spmm(self.h, n_fields, d_u.ptr, d_out.ptr, field_type, accum), "spmm")

    def tune_set(self, kind, n_fields, field_type, accum, fpt, block, batch=4, share=0):
        self.ctx.check(RBFX.rbfx_tune_set(self.h, kind, n_fields, field_type, accum, fpt, block, batch, share), "tune_set")

    def tune_get(self, kind, n_fields, field_type, accum):
        """(fields per thread, block, batch, share) the library will launch with."""
        f, b, u, sh = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        self.ctx.check(RBFX.rbfx_tune_get(self.h, kind, n_fields, field_type, accum, C.byref(f), C.byref(b), C.byref(u),
                                          C.byref(sh)), "tune_get")
        return f.value, b.value, u.value, sh.value

    def set_row_order(self, order):
        """Process rows in `order` (a permutation; host array or device array); results unchanged."""
        d = order if isinstance(order, (DeviceArray, DeviceView)) else self.ctx.to_device(np.asarray(order, np.int32))
        self.ctx.check(RBFX.rbfx_ell_set_row_order(self.h, d.ptr), "ell_set_row_order")
        self.ctx.sync()

    def build_local(self, tile_rows=64, max_nodes=0):
        """Tables for LDS-staged gathers (tiles of at most tile_rows positions and max_nodes
        distinct nodes); returns (tile_rows, k_max, n_tiles)."""
        self.ctx.check(RBFX.rbfx_ell_build_local(self.h, tile_rows, max_nodes), "ell_build_local")
        return self.local_info()

    def local_info(self):
        t, k, n = C.c_int32(), C.c_int32(), C.c_int32()
        self.ctx.check(RBFX.rbfx_ell_local_info(self.h, C.byref(t), C.byref(k), C.byref(n)), "ell_local_info")
        return t.value, k.value, n.value

    def close(self):
        if self.h:
            RBFX.rbfx_ell_destroy(self.h)
            self.h = None


class RK4:
    def __init__(self, ctx, ell, n_fields, field_storage=F64, accum=F64, coefs=None, use_graph=True):
        self.ctx = ctx
        d = _RK4Desc()
        d.ell = ell.h
        d.n_fields = n_fields
        d.field_storage = field_storage
        d.accum = accum
        self._coefs = []
        for o in range(ell.n_ops):
            c = None if coefs is None else coefs[o]
            if c is not None:
                dc = c if isinstance(c, (DeviceArray, DeviceView)) else ctx.to_device(np.asarray(c, np.float64))
                self._coefs.append(dc)
                d.d_coef[o] = dc.ptr.value
        d.use_graph = 1 if use_graph else 0
        h = C.c_void_p()
        ctx.check(RBFX.rbfx_rk4_create(ctx.h, C.byref(d), C.byref(h)), "rk4_create")
        self.h = h
        self.n_rows = ell.n_rows
        self.n_fields = n_fields

    def set_state(self, u):
        d = u if isinstance(u, (DeviceArray, DeviceView)) else self.ctx.to_device(np.asarray(u, np.float64))
        self.ctx.check(RBFX.rbfx_rk4_set_state(self.h, d.ptr), "rk4_set_state")
        self._state_in = d

    def get_state(self):
        d = self.ctx.alloc((self.n_rows, self.n_fields), np.float64)
        self.ctx.check(RBFX.rbfx_rk4_get_state(self.h, d.ptr), "rk4_get_state")
        out = d.download()
        d.free()
        return out

    def step(self, dt, n_steps=1):
        self.ctx.check(RBFX.rbfx_rk4_step(self.h, dt, n_steps), "rk4_step")

    def stage(self, stage, dt):
        self.ctx.check(RBFX.rbfx_rk4_stage(self.h, stage, dt), "rk4_stage")

    def close(self):
        if self.h:
            RBFX.rbfx_rk4_destroy(self.h)
            self.h = None


class Stokes:
    """Steady Stokes system on the sphere (rbfx_stokes_*): ops64 / ops32 are EllOperators with the
    LSFC, XSFC, YSFC, ZSFC weights in FP64 / FP32 storage; vectors are 4N + 4 doubles in the
    block layout u, v, w, p, constraint multipliers."""

    def __init__(self, ctx, ops64, ops32=None, eta=1.0):
        self.ctx = ctx
        self._ops = (ops64, ops32)  # the handle reads their device arrays
        h = C.c_void_p()
        ctx.check(RBFX.rbfx_stokes_create(ctx.h, ops64.h, ops32.h if ops32 is not None else None, eta, C.byref(h)),
                  "stokes_create")
        self.h = h
        n, nl, nu = C.c_int32(), C.c_int32(), C.c_int32()
        ctx.check(RBFX.rbfx_stokes_info(h, C.byref(n), C.byref(nl), C.byref(nu)), "stokes_info")
        self.n, self.lower_levels, self.upper_levels = n.value, nl.value, nu.value
        self.N = ops64.n_rows
        self.width = ops64.width

    def _dev(self, v):
        return v if isinstance(v, (DeviceArray, DeviceView)) else self.ctx.to_device(np.asarray(v, np.float64))

    def apply(self, x):
        dx = self._dev(x)
        dy = self.ctx.alloc(self.n, np.float64)
        self.ctx.check(RBFX.rbfx_stokes_apply(self.h, dx.ptr, dy.ptr), "stokes_apply")
        out = dy.download()
        dy.free()
        return out

    def factor(self):
        d = self.ctx.alloc((self.N, self.width), np.float64)
        self.ctx.check(RBFX.rbfx_stokes_factor(self.h, d.ptr), "stokes_factor")
        out = d.download()
        d.free()
        return out

    def precond(self, v):
        d = self.ctx.to_device(np.array(v, np.float64))
        self.ctx.check(RBFX.rbfx_stokes_precond(self.h, d.ptr), "stokes_precond")
        out = d.download()
        d.free()
        return out

    def gmres(self, b, tol=1e-8, max_iter=300, krylov=50, mode=F64, device_out=False):
        """Returns (x, iterations, error); x stays on the device with device_out."""
        db = self._dev(b)
        dx = self.ctx.alloc(self.n, np.float64)
        it, err = C.c_int32(), C.c_double()
        self.ctx.check(RBFX.rbfx_stokes_gmres(self.h, db.ptr, dx.ptr, tol, max_iter, krylov, mode, C.byref(it),
                                              C.byref(err)), "stokes_gmres")
        if device_out:
            return dx, it.value, err.value
        out = dx.download()
        dx.free()
        return out, it.value, err.value

    def close(self):
        if self.h:
            RBFX.rbfx_stokes_destroy(self.h)
            self.h = None
