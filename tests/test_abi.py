"""The C-ABI library loads without a GPU and exports every entry point include/*.h declares;
host-only entry points work; device entry points fail cleanly (no crash) without a device."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import rbffd_gpu_amd as R
from conftest import ROOT


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(inc, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(rbfx_[a-z0-9_]+)\s*\(", src, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("rbfx_create", "rbfx_weights", "rbfx_ell_create", "rbfx_spmv", "rbfx_spmm", "rbfx_rk4_create",
                 "rbfx_rk4_step", "rbfx_knn_kdtree"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(R.lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", R.lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (rbfx_[a-z0-9_]+)", out))
    assert set(declared_functions()) <= exported


def test_host_entry_points_without_gpu():
    x = R.nodes_fibonacci(100, 0.0, 0)
    assert np.allclose(np.linalg.norm(x, axis=1), 1.0)
    st = R.knn_kdtree(x, 8)
    assert st.shape == (100, 8)


def test_invalid_arguments_are_reported_not_fatal():
    lib = R.lib.load_library()
    assert lib.rbfx_knn_kdtree(0, None, 5, None, 0) == -1
    assert lib.rbfx_nodes_fibonacci(-3, 0.0, 0, None) == -1
    n = C.c_int32()
    assert lib.rbfx_nodes_icosahedral(11, None, C.byref(n)) == -1
    assert lib.rbfx_sync(None) == -1
    assert lib.rbfx_knn_gpu(None, 10, None, 3, None, None) == -1
    assert lib.rbfx_knn_gpu_device(None, 10, None, 3, None, None) == -1
    assert lib.rbfx_weights(None, 1, None, None, 1, None, None, 1, 3, 0, 0, 1.0, 0, None, None, None) == -1


def test_create_without_device_fails_cleanly():
    """On a host without a GPU, rbfx_create returns RBFX_E_NODEVICE (or succeeds on a GPU host)."""
    lib = R.lib.load_library()
    h = C.c_void_p()
    rc = lib.rbfx_create(0, C.byref(h))
    if rc == 0:
        lib.rbfx_destroy(h)
    else:
        assert rc in (-4, -2)
        assert lib.rbfx_last_error(None)


def test_cpp_layer_builds_and_exports():
    """The drop-in C++ classes are compiled into librbfx_cxx.so next to librbfx.so."""
    p = os.path.join(ROOT, "rbffd_gpu_amd", "librbfx_cxx.so")
    assert os.path.exists(p), "build the C++ layer: make -C rbffd_gpu_amd cxx"
    out = subprocess.run(["nm", "-DC", "--defined-only", p], capture_output=True, text=True).stdout
    for sym in ("rbfx::RBFFD_HIP::applyWeightsForDeriv", "rbfx::RBFFD_HIP::computeAllWeightsForAllStencils",
                "rbfx::Grid::generateStencils", "rbfx::TimeDependentPDE_HIP::advance", "rbfx::CosineBell_HIP::CosineBell_HIP"):
        assert sym in out, sym
