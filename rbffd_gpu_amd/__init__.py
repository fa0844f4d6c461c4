"""rbffd_gpu_amd: MI355X (gfx950) RBF-FD hot path behind a C ABI (include/rbfx.h).

This Python module is plumbing for tests and benchmarks: it loads the in-tree librbfx.so with
ctypes and wraps its C entry points. It has no compute of its own and no CPU fallback: if
the library is missing, importing it raises.
"""
from .lib import (  # noqa: F401
    RBFX,
    RbfxError,
    Context,
    DeviceArray,
    EllOperators,
    RK4,
    Stokes,
    load_library,
    dertype_bits,
    X, Y, Z, LAPL, R, HV, LAMBDA, THETA, LSFC, XSFC, YSFC, ZSFC, INTERP,
    F64, F32, F32_RN,
    SOLVE_F64_LU, SOLVE_F32_LU_F64_IR, POLY_SPHERE,
    knn_kdtree,
    knn_gpu,
    knn_gpu_device,
    nodes_fibonacci,
    nodes_icosahedral,
    order_morton,
    sph_harm,
    make_fields,
    csr_from_fixed,
)
