import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(autouse=True)
def _fresh_tuning_table():
    """Each test starts and ends with an empty launch-shape table (it is process-global)."""
    import rbffd_gpu_amd as R

    yield
    try:
        R.lib.RBFX.rbfx_tune_clear()
    except OSError:
        pass


@pytest.fixture(scope="session")
def ctx():
    import rbffd_gpu_amd as R

    c = R.Context(0)
    print("device:", c.info(), flush=True)
    yield c
    c.close()


@pytest.fixture(scope="session")
def c1():
    """Config-1 inputs: committed node fixture, reference kd-tree stencils (n=31)."""
    d = np.load(os.path.join(GOLDEN, "c1_knn.npz"))
    return {"xyz": d["xyz"], "st": d["st_ref"]}


def rel_apply_err(y, y_ref, absum):
    """|y - y_ref| normalised by sum_j |w_ij u_j| (cancellation-aware parity metric)."""
    return np.max(np.abs(y - y_ref) / np.maximum(absum, 1e-300))
