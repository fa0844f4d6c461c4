"""GPU k-nearest-neighbour stencils against the host kd-tree (itself bit-identical to the
reference's external/kdtree search, tests/test_knn.py): index lists must match bit for bit.

Node sets cover the cases that decide the split between the GPU search and the host
fallback: jittered spiral nodes (distinct distances almost everywhere), the icosahedral grid
(symmetric, many exactly equal distances), a volume and a plane, duplicated nodes (distance
ties at 0), and k close to N.
"""
import numpy as np
import pytest

import rbffd_gpu_amd as R

pytestmark = pytest.mark.gpu


def _ico(level):
    import ctypes as C
    n = C.c_int32()
    R.RBFX.rbfx_nodes_icosahedral(level, None, C.byref(n))
    xyz = np.empty((n.value, 3))
    assert R.RBFX.rbfx_nodes_icosahedral(level, xyz.ctypes.data_as(C.c_void_p), C.byref(n)) == 0
    return xyz


def _sets():
    rng = np.random.default_rng(4)
    plane = rng.random((3000, 3))
    plane[:, 2] = 0.5
    dup = R.nodes_fibonacci(2000, 1e-3, 2)
    dup[1500:1520] = dup[100:120]  # 20 duplicated nodes: ties at distance 0
    return [("fibonacci 20000", R.nodes_fibonacci(20000, 1e-3, 0), (31, 50)),
            ("icosahedral level 4", _ico(4), (31, 50)),
            ("cube 5000", rng.random((5000, 3)), (17, 50)),
            ("plane 3000", plane, (31,)),
            ("duplicates 2000", dup, (17,)),
            ("small 60", R.nodes_fibonacci(60, 1e-3, 1), (59, 60))]


@pytest.mark.parametrize("case", range(6))
def test_knn_gpu_matches_kdtree(ctx, case):
    name, xyz, ks = _sets()[case]
    for k in ks:
        print(f"case: knn_gpu {name} k={k}", flush=True)
        ref = R.knn_kdtree(xyz, k)
        got, n_host = R.knn_gpu(ctx, xyz, k)
        print(f"  {n_host} of {xyz.shape[0]} rows from the host kd-tree")
        assert np.array_equal(got, ref)
        if not name.startswith("duplicates"):
            assert np.array_equal(got[:, 0], np.arange(xyz.shape[0]))  # each row starts with its node
        if name.startswith("fibonacci"):
            assert n_host < xyz.shape[0] // 100  # distinct distances: the GPU answers nearly every row
        if name.startswith("duplicates"):
            assert n_host >= 40  # both copies of each duplicated node tie at distance 0


def test_knn_gpu_rejects_k_above_n(ctx):
    with pytest.raises(R.RbfxError):
        R.knn_gpu(ctx, R.nodes_fibonacci(10, 1e-3, 0), 11)


def test_knn_gpu_reference_golden_lists(ctx, c1):
    """Against the reference's own kd-tree output (fixtures from oracle/_ref/ref_knn): the C1
    node set and the tie-heavy grid, lattice and duplicate sets, bit for bit."""
    import os
    from conftest import GOLDEN
    got, _ = R.knn_gpu(ctx, c1["xyz"], 31)
    assert np.array_equal(got, c1["st"])
    g = np.load(os.path.join(GOLDEN, "ties_knn.npz"))
    for name, ks in (("grid", (5, 9, 13, 31)), ("lattice", (7, 27, 50)), ("dups", (9, 17))):
        for k in ks:
            got, n_host = R.knn_gpu(ctx, g[name], k)
            print(f"  {name} k={k}: {n_host} of {g[name].shape[0]} rows from the host kd-tree")
            assert np.array_equal(got, g[f"{name}_k{k}"]), (name, k)


def test_knn_gpu_large_sets_digest(ctx):
    """N = 163842 with k = 101 (32 queries per block) and N = 10^6 with k = 50: SHA-256 of the
    GPU stencil array equals that of the reference's kd-tree output."""
    import hashlib
    import json
    import os
    from conftest import GOLDEN
    with open(os.path.join(GOLDEN, "large_knn.json")) as f:
        big = json.load(f)
    for key, rec in big.items():
        print(f"case: knn_gpu digest {key}", flush=True)
        xyz = R.nodes_fibonacci(rec["n"], rec["jitter"], rec["seed"])
        assert hashlib.sha256(xyz.tobytes()).hexdigest() == rec["xyz_sha256"], "node generator changed"
        st, n_host = R.knn_gpu(ctx, xyz, rec["k"])
        print(f"  {n_host} of {rec['n']} rows from the host kd-tree")
        assert np.array_equal(st[:16], np.array(rec["first_rows"], np.int32))
        assert hashlib.sha256(st.tobytes()).hexdigest() == rec["st_sha256"], key


def test_knn_gpu_device_output(ctx):
    """rbfx_knn_gpu_device leaves the same lists in device memory: tie-free rows from the GPU
    search, tie rows from the host kd-tree uploaded row by row, and k = N entirely from the
    host kd-tree."""
    cases = (("fibonacci 20000", R.nodes_fibonacci(20000, 1e-3, 0), 50),
             ("icosahedral level 4", _ico(4), 31),
             ("small 60", R.nodes_fibonacci(60, 1e-3, 1), 60))
    for name, xyz, k in cases:
        print(f"case: knn_gpu_device {name} k={k}", flush=True)
        d_st, n_host = R.knn_gpu_device(ctx, xyz, k)
        got = d_st.download()
        d_st.free()
        print(f"  {n_host} of {xyz.shape[0]} rows from the host kd-tree")
        assert np.array_equal(got, R.knn_kdtree(xyz, k)), name
