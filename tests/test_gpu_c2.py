"""The cosine-bell configuration at its full size, through the C++ drop-in layer's driver
(rbfx_cosine_bell: Grid -> RBFFD_HIP -> CosineBell_HIP, the reference's
tests/experiments/cosine_bell_vcl/main.cpp:495-585 flow): icosahedral level 7 (N = 163,842
nodes on the unit sphere), n = 101 kd-tree stencils, Gaussian RBF with the constant monomial
(the reference's ONE_MONOMIAL), LAMBDA / THETA / Delta^4 hyperviscosity (dim 2,
gamma = 2^-6), flow angle pi/2, one revolution in 8000 RK4 steps.

eps = 0.065 sqrt(N). The specified 0.035 sqrt(N) puts these 101-node systems at kappa ~ 1e19
(far into the flat limit): the solve returns weights with a small backward error that are
meaningless, and advection with them blows up in FP64 for every gamma and step count tried;
0.065 is the most accurate stable value found (kappa ~ 1e13-1e14).

Checks: the GPU weights' normwise backward error on a sample of stencils (the oracle's
assembly of the same systems), their forward agreement with the oracle's weights within
64 kappa u; 20 RK4 steps of the driver replayed on the CPU with the GPU weights (1e-12); one
revolution in FP64 and in FP32 storage: the error against the exact bell (which after one
revolution equals the initial bell) and the FP32-vs-FP64 deviation against stated bounds."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import rbffd_gpu_amd as R
from conftest import ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "rbffd_gpu_amd", "bin", "rbfx_cosine_bell")
U64 = 2.220446049250313e-16
LEVEL, NST, C1, GAMMA, STEPS = 7, 101, 0.065, 2.0 ** -6, 8000
T_REV = 1036800.0


def run(args, cwd):
    print("run:", " ".join(args), flush=True)
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return json.loads(p.stdout.strip().splitlines()[-1])


def driver(prec, tmp_path, max_steps=0, dump=0):
    return run([BIN, f"ico:{LEVEL}", str(NST), str(C1), "0", "4", repr(GAMMA), str(STEPS), "1", prec, "2", "2",
                "morton", "kdtree", str(max_steps), str(dump)], tmp_path)


@pytest.fixture(scope="module")
def c2_nodes():
    xyz = R.nodes_icosahedral(LEVEL)
    xyz = np.ascontiguousarray(xyz[R.order_morton(xyz)])  # the driver's Morton numbering
    st = R.knn_kdtree(xyz, NST)
    return xyz, st


def test_c2_weights_and_20_steps(tmp_path, c2_nodes):
    xyz, st = c2_nodes
    N = xyz.shape[0]
    assert N == 163842
    res = driver("f64", tmp_path, max_steps=20, dump=1)
    print("driver:", res, flush=True)
    m = NST + 1
    assert res["max_backward_error"] <= 64 * m * U64
    w_gpu = np.fromfile(tmp_path / "cosine_bell_weights.bin", dtype=np.float64).reshape(3, N, NST)  # HV, LAMBDA, THETA
    eps = C1 * np.sqrt(N)
    hv_scale = -GAMMA * N ** -4.0
    mask = R.LAMBDA | R.THETA | R.HV  # columns in ascending bit order: HV, LAMBDA, THETA (as in the file)
    rows = np.arange(0, N, N // 48)
    off = np.arange(rows.size + 1, dtype=np.int32) * NST
    idx = np.ascontiguousarray(st[rows]).astype(np.int32).ravel()
    w_ref, _, ns = O.weights(off, idx, xyz, eps, mask, dim=2, poly_deg=0, hv_k=4, hv_scale=hv_scale)
    assert ns == 0
    worst_be, worst_ratio = 0.0, 0.0
    for q, s in enumerate(rows):
        M, Rr = O.stencil_system(xyz, st[s], eps, mask, 2, 0, 4)
        kappa = np.linalg.cond(M)
        X = np.zeros((m, 3))
        for t in range(3):
            w = w_gpu[t, s]
            X[:NST, t] = w / (hv_scale if t == 0 else 1.0)
        # the constant monomial's multiplier is not stored: take the one that minimises the
        # constraint-free rows' residual (the last row of M W = R holds sum_j w_j = 0)
        X[NST, :] = np.linalg.lstsq(M[:NST, NST:], (Rr - M[:, :NST] @ X[:NST])[:NST], rcond=None)[0]
        r = Rr - M @ X
        be = np.abs(r).max() / (np.abs(M).sum(1).max() * np.abs(X).max() + np.abs(Rr).max())
        worst_be = max(worst_be, be)
        for t in range(3):
            ref = w_ref[t, q * NST:(q + 1) * NST]
            rel = np.abs(w_gpu[t, s] - ref).max() / np.abs(ref).max()
            assert rel <= min(1.0, 64 * kappa * U64), (s, t, rel, kappa)
            worst_ratio = max(worst_ratio, rel / (kappa * U64))
    print(f"sample of {rows.size} stencils: backward error max {worst_be:.2e}, forward diff / (kappa u) max "
          f"{worst_ratio:.2e}", flush=True)
    assert worst_be <= 64 * m * U64
    # 20 RK4 steps replayed on the CPU with the GPU weights
    u_gpu = np.fromfile(tmp_path / "cosine_bell_solution.bin", dtype=np.float64)
    cl, ct = O.cosine_bell_coefs(xyz)
    h0 = O.cosine_bell_exact(xyz)
    u = O.rk4_step(st, [w_gpu[1], w_gpu[2], w_gpu[0]], [cl, ct, None], h0[:, None], T_REV / STEPS, n_steps=20)[:, 0]
    err = np.abs(u_gpu - u).max() / np.abs(u).max()
    print(f"20 steps: GPU vs CPU replay {err:.2e}", flush=True)
    assert err <= 1e-12


def test_c2_one_revolution_fp64_and_fp32(tmp_path):
    d64, d32 = tmp_path / "f64", tmp_path / "f32"
    d64.mkdir()
    d32.mkdir()
    r64 = driver("f64", d64)
    r32 = driver("f32", d32)
    u64 = np.fromfile(d64 / "cosine_bell_solution.bin", dtype=np.float64)
    u32 = np.fromfile(d32 / "cosine_bell_solution.bin", dtype=np.float64)
    dev = np.abs(u32 - u64).max() / np.abs(u64).max()
    print(f"one revolution: FP64 l2rel {r64['l2rel']:.4e} linfrel {r64['linfrel']:.4e} ({r64['steps_per_s']} steps/s); "
          f"FP32 l2rel {r32['l2rel']:.4e} linfrel {r32['linfrel']:.4e} ({r32['steps_per_s']} steps/s); "
          f"FP32 vs FP64 max deviation {dev:.2e}", flush=True)
    for r in (r64, r32):  # measured 4.42e-4 and 5.04e-4 in both precisions
        assert r["l2rel"] <= 1e-3 and r["linfrel"] <= 1e-3
    assert dev <= 1e-6  # measured 2.7e-7
    assert abs(r32["l2rel"] - r64["l2rel"]) <= 1e-3 * r64["l2rel"]
