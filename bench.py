"""Benchmark: RK4 steps/s and SpMV HBM GB/s at N=1M, n=50 on one MI355X (FP32-stored vs FP64).

Workload (the fused multi-operator multi-field RK4 configuration):
  N = 1,000,000 nodes on the unit sphere (golden-angle spiral, +-1e-3 jitter, seed 0,
  projected back onto the sphere), n = 50 kd-tree stencils, Gaussian RBF
  eps = 4 (0.044 sqrt(N) - 0.14) (the reference's n=50 shape parameter is too flat for
  stable advection on spiral nodes), degree-2 polynomials in the 9-function basis that is
  independent on the sphere,
  4 operators: d/dx, d/dy, d/dz on the sphere (the reference's XSFC/YSFC/ZSFC, the Cartesian
  gradient projected onto the tangent plane; unprojected d/dx on nodes that all lie on the
  sphere leaves the normal direction undetermined and its weights blow up) and
  hyperviscosity Delta^4 (dim 2) scaled -gamma N^-4 with gamma = 2^9 per revolution (the
  run is stable over a revolution for gamma = 2^9 .. 2^10; with 2^8 or less a spurious
  growing mode of the advection operator blows the FP64 solution up within 1000 steps, from
  2^11 on the hyperviscosity term leaves the RK4 stability region at this dt),
  16 fields = seeded combinations of real spherical harmonics of degree <= 20,
  PDE: solid-body rotation advection dq/dt = -(w x r).grad_S q + H q (one revolution per
  unit time), classic RK4, dt = 1/4000.
One "step" = one RK4 step of all 16 fields (4 fused stage kernels, replayed from a hipGraph)
with every input resident in HBM. Stencils and weights are built on the device before the
timed region (weights: batched GPU solve).

Printed JSON: value = FP32-stored RK4 steps/s (FP32 weights and stage fields, FP32 gather
accumulation as in the reference's float mode, FP64 state and RK4 combination); extras carry
the FP64 path, the FP32-stored path with FP64 accumulation, the SpMV GB/s, the FP32-vs-FP64
deviation and the error against the analytic rotated field.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rotate(xyz, axis, angle):
    k = axis / np.linalg.norm(axis)
    c, s = np.cos(angle), np.sin(angle)
    return xyz * c + np.cross(k, xyz) * s + np.outer(xyz @ k, k) * (1 - c)


def bytes_per_stage(N, n, ops, F, s_w, s_f, n_coef, stage):
    """Compulsory HBM bytes of one fused RK4 stage kernel launch:
    weights + indices once, the gathered stage field once per node, the FP64 state u,
    the FP64 accumulator (written in stage 1, read+written in 2-3, read in 4), the next
    stage field (stage 4 of the FP64 path writes u only), the coefficient arrays."""
    w = ops * N * n * s_w + N * n * 4
    coef = n_coef * N * 8
    f64 = s_f == 8
    if stage == 1:
        fld = s_f + 8 + 8 + s_f  # gather X (for FP64 that is u itself), u, acc write, X' write
        if f64:
            fld = 8 + 8 + 8  # u read once serves gather and state
    elif stage in (2, 3):
        fld = s_f + 8 + 16 + s_f
    else:
        fld = s_f + 8 + 8 + 8 + (0 if f64 else s_f)
    return w + coef + F * N * fld


STAGE_PROFILE = os.path.join(ROOT, "profiles", "rk4_stage_N1M_n50_ops4_F16.md")


def profile_traffic(kernel):
    """HBM bytes per launch of `kernel` measured by rocprofv3 FETCH_SIZE / WRITE_SIZE passes over
    this same workload (scaled by the copy-kernel calibration), read from the committed profile
    summary's traffic table: (bytes, profiled kernel). The exact instance is preferred; else the
    same kernel and element types with another tuned launch shape (the autotuner may pick
    another shape from run to run; the bytes moved differ little between shapes). (None, None)
    when the profile holds no such row."""
    rows = {}
    try:
        with open(STAGE_PROFILE) as fh:
            for line in fh:
                cells = [c.strip() for c in line.split("|")]
                # traffic table rows: kernel | fetched | written | traffic | algorithmic | ratio | us | GB/s
                if len(cells) == 10 and cells[1].startswith("`") and cells[4][:1].isdigit():
                    rows[cells[1].strip("`")] = float(cells[4])
    except OSError:
        return None, None
    if kernel in rows:
        return rows[kernel], kernel
    family = kernel.split(",")[:3]
    for name, b in rows.items():
        if name.split(",")[:3] == family:
            return b, name
    return None, None


def cpu_model():
    """Host CPU model name (first 'model name' line of /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


_ORACLE = None


def cpu_oracle():
    """The CPU restatement for the baseline legs, compiled -O3 -march=native (OpenMP,
    -ffp-contract=off) for the host that runs the bench; the prebuilt x86-64-v3 build if the
    host cannot compile. Returns (module, description of the build)."""
    global _ORACLE
    if _ORACLE is None:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import tempfile

        import oracle as O

        d = tempfile.mkdtemp(prefix="rbfx_oracle_")
        path = O.build_native(d)
        if path:
            O.use_library(path)
            _ORACLE = (O, "-O3 -march=native -fopenmp -ffp-contract=off")
        else:
            _ORACLE = (O, "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off (native build failed)")
    return _ORACLE


def median_rate(fn, runs=5, warmup=1):
    """fn() -> (units, seconds); warm-up calls untimed, then the median rate of `runs` calls."""
    for _ in range(warmup):
        fn()
    rates = []
    for _ in range(runs):
        units, sec = fn()
        rates.append(units / sec)
    return float(np.median(rates)), rates


def stream_copy_gbs(ctx, R, nbytes=1 << 30, reps=10):
    """Measured HBM bandwidth of a plain copy kernel (read + write bytes / time)."""
    a = ctx.alloc(nbytes // 4, np.float32)
    b = ctx.alloc(nbytes // 4, np.float32)
    ctx.check(R.RBFX.rbfx_stream_copy(ctx.h, b.ptr, a.ptr, nbytes))
    ctx.timer_start()
    for _ in range(reps):
        ctx.check(R.RBFX.rbfx_stream_copy(ctx.h, b.ptr, a.ptr, nbytes))
    ms = ctx.timer_stop() / reps
    a.free()
    b.free()
    return 2 * nbytes / (ms * 1e-3) / 1e9


C3_EPS = {17: (0.026, 0.08), 31: (0.035, 0.1), 50: (0.044, 0.14), 75: (0.051, 0.15), 101: (0.058, 0.16)}


KAPPA_EDGES = (6, 8, 10, 12, 14)  # log10 bin edges of the 2-norm condition number


def kappa_bins(xyz, off, idx, eps, mask, deg, runs, n_sample=1500, seed=3):
    """Weight error by condition number on a seeded sample of the stencils: each stencil's
    system is assembled by the CPU restatement, kappa = cond_2(M) (numpy SVD), and its
    reference weights are the restatement's FP64 LU. Per bin and solve mode: stencil count,
    max normwise backward error (the GPU diagnostics), max forward difference
    max|w - w_ref| / max|w_ref|, that difference over kappa * u, and the FP64 fallbacks of the
    FP32 + refinement mode."""
    O, _ = cpu_oracle()
    u64 = np.finfo(np.float64).eps / 2
    rng = np.random.default_rng(seed)
    sample = np.sort(rng.choice(off.size - 1, size=n_sample, replace=False))
    s_off = np.concatenate([[0], np.cumsum(np.diff(off)[sample])]).astype(np.int32)
    s_idx = np.concatenate([idx[off[s]:off[s + 1]] for s in sample]).astype(np.int32)
    w_ref, _, _ = O.weights(s_off, s_idx, xyz, eps[sample], mask, dim=3, poly_deg=deg)
    kap = np.array([np.linalg.cond(O.stencil_system(xyz, idx[off[s]:off[s + 1]], eps[s], mask, 3, deg, 0)[0])
                    for s in sample])
    lk = np.log10(kap)
    names = ["<1e6"] + [f"1e{a}-1e{b}" for a, b in zip(KAPPA_EDGES[:-1], KAPPA_EDGES[1:])] + [f">=1e{KAPPA_EDGES[-1]}"]
    which = np.searchsorted(np.array(KAPPA_EDGES, float), lk, side="right")
    out = {}
    for b, name in enumerate(names):
        sel = np.nonzero(which == b)[0]
        if sel.size == 0:
            continue
        entry = {"stencils": int(sel.size), "kappa_max": float(kap[sel].max())}
        for mode, (w, dg) in runs.items():
            fwd, ratio = [], []
            for q in sel:
                s = sample[q]
                ref = w_ref[:, s_off[q]:s_off[q + 1]]
                d = np.abs(w[:, off[s]:off[s + 1]] - ref).max() / np.abs(ref).max()
                fwd.append(d)
                ratio.append(d / (kap[q] * u64))
            e = {"max_backward_error": float(np.nanmax(dg[sample[sel], 0])), "max_forward_diff": float(max(fwd)),
                 "max_forward_diff_over_kappa_u": float(max(ratio))}
            if mode == "f32_lu_f64_ir":
                e["fp64_fallbacks"] = int((dg[sample[sel], 3].astype(int) == 2).sum())
            entry[mode] = e
        out[name] = entry
    return out


def bench_weights(ctx, R, N, cpu, nthreads, degrees=(2, 3), modes=("f64_lu", "f32_lu_f64_ir"), eps_scale=1.0,
                  size_set=(17, 31, 50, 75, 101)):
    """Batched weights with uneven stencils: N jittered sphere nodes, each stencil the n nearest
    with n drawn uniformly from {17, 31, 50, 75, 101} (seed 1), eps = c1 sqrt(N) - c2 per size
    (the reference's table, n = 75 interpolated), X/Y/Z/LAPL (dim 3), degree 2 and 3; FP64 LU and
    FP32 LU + FP64 refinement, each one timed call over all stencils without diagnostics (the
    production path, after one untimed warm-up call), then an untimed call with diagnostics for
    the backward errors. With cpu, the C++ restatement
    times a bounded sample of the same stencils on the host threads and checks the GPU weights
    of that sample (backward error, forward difference against the per-stencil condition).
    eps_scale multiplies every eps (1: the reference's table; larger: less flat, better
    conditioned systems). size_set restricts the sizes drawn (one size: one class, for per-class
    timing)."""
    t0 = time.time()
    xyz = R.nodes_fibonacci(N, 1e-3, 0)
    knn = R.knn_kdtree(xyz, 101)
    sizes = np.random.default_rng(1).choice(list(size_set), size=N)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    mask_cols = np.arange(101)[None, :] < sizes[:, None]
    idx = knn[mask_cols].astype(np.int32)
    eps = np.array([C3_EPS[k][0] * np.sqrt(N) - C3_EPS[k][1] for k in (17, 31, 50, 75, 101)])[
        np.searchsorted([17, 31, 50, 75, 101], sizes)] * eps_scale
    log(f"weights workload: {N} stencils, {int(off[-1])} nonzeros, setup {time.time() - t0:.1f}s")
    d_xyz, d_off, d_idx = ctx.to_device(xyz), ctx.to_device(off), ctx.to_device(idx)
    mask = R.X | R.Y | R.Z | R.LAPL
    res = {"N": N, "sizes": [int(k) for k in size_set], "operators": "X,Y,Z,LAPL (dim 3)", "eps_scale": eps_scale}
    outs = {}
    all_modes = {"f64_lu": R.SOLVE_F64_LU, "f32_lu_f64_ir": R.SOLVE_F32_LU_F64_IR}
    for deg in degrees:
        dgs = {}
        for mode_name, mode in ((k, all_modes[k]) for k in modes):
            # one untimed call first (loads the kernels' code objects), then the timed call
            d_w, _ = ctx.weights(d_off, d_idx, d_xyz, eps, mask, dim=3, poly_deg=deg, mode=mode, device_out=True)
            d_w.free()
            ctx.sync()
            t = time.perf_counter()
            d_w, _ = ctx.weights(d_off, d_idx, d_xyz, eps, mask, dim=3, poly_deg=deg, mode=mode, device_out=True)
            ctx.sync()
            sec = time.perf_counter() - t
            d_w.free()
            # diagnostics (backward error, path) from a second, untimed call of the same solve
            d_w, d_dg = ctx.weights(d_off, d_idx, d_xyz, eps, mask, dim=3, poly_deg=deg, mode=mode, diag=True,
                                    device_out=True)
            dg = d_dg.download()
            d_dg.free()
            dgs[mode_name] = dg
            paths = dg[:, 3].astype(int)
            entry = {"s": round(sec, 4), "stencils_per_s": round(N / sec, 1),
                     "max_backward_error": float(np.nanmax(dg[:, 0])),
                     "fp64_fallbacks": int((paths == R.lib.PATH_F32_FALLBACK_F64).sum())}
            outs[(deg, mode_name)] = d_w
            res[f"deg{deg}_{mode_name}"] = entry
            log(f"weights deg {deg} {mode_name}: {sec:.3f}s ({N / sec:.3g} stencils/s), {entry}")
        if len(modes) == 2:  # FP32 + refinement against the FP64 LU weights (all stencils)
            w64 = outs[(deg, "f64_lu")].download()
            w32 = outs[(deg, "f32_lu_f64_ir")].download()
            scale = np.abs(w64).max(axis=0)
            res[f"deg{deg}_ir_vs_f64_max_rel"] = float((np.abs(w32 - w64).max(axis=0) /
                                                        np.maximum(scale, 1e-300)).max())
            if cpu:
                res[f"deg{deg}_kappa_bins"] = kappa_bins(xyz, off, idx, eps, mask, deg,
                                                         {"f64_lu": (w64, dgs["f64_lu"]),
                                                          "f32_lu_f64_ir": (w32, dgs["f32_lu_f64_ir"])})
                log(f"weights deg {deg} error by condition number: {res[f'deg{deg}_kappa_bins']}")
        for key in list(outs):
            outs[key].free()
        outs.clear()
    if cpu:
        O, build = cpu_oracle()
        rng = np.random.default_rng(2)
        sample = np.sort(rng.choice(N, size=100000, replace=False))
        s_off = np.concatenate([[0], np.cumsum(sizes[sample])]).astype(np.int32)
        s_idx = np.concatenate([idx[off[s]:off[s + 1]] for s in sample]).astype(np.int32)
        w_cpu, dg_cpu, _ = O.weights(s_off, s_idx, xyz, eps[sample], mask, dim=3, poly_deg=2, diag=True,
                                     n_threads=nthreads)

        def leg(ns, threads):
            def f():
                t = time.perf_counter()
                O.weights(s_off[:ns + 1], s_idx[:s_off[ns]], xyz, eps[sample[:ns]], mask, dim=3, poly_deg=2,
                          n_threads=threads)
                return ns, time.perf_counter() - t
            return f

        rate, rates = median_rate(leg(len(sample), nthreads), runs=5, warmup=1)
        w_gpu, _ = ctx.weights(s_off, s_idx, xyz, eps[sample], mask, dim=3, poly_deg=2, diag=True)
        rel = np.abs(w_gpu - w_cpu).max() / np.abs(w_cpu).max()
        res["cpu_baseline"] = {"stencils_per_s": round(rate, 1), "cores": nthreads, "kind": "port",
                               "sample": f"{len(sample)} stencils of the same size mix, degree 2, FP64 LU "
                                         f"(oracle/oracle.cpp, {build}), OpenMP over stencils; median of 5 timed "
                                         f"runs after one warm-up run",
                               "runs_stencils_per_s": [round(r, 1) for r in rates],
                               "gpu_vs_cpu_max_rel_weight_diff": float(rel),
                               "cpu_max_backward_error": float(dg_cpu[:, 0].max())}
        # the same solves on one host thread, on the first 10^4 stencils of the sample
        s1 = 10000
        r1, _ = median_rate(leg(s1, 1), runs=3, warmup=0)
        res["cpu_baseline"]["one_thread"] = {"stencils_per_s": round(r1, 1),
                                             "sample": f"{s1} stencils, median of 3 runs"}
        res["cpu_baseline"]["cpu_model"] = cpu_model()
        res["cpu_baseline"]["nproc"] = os.cpu_count()
        res["gpu_over_cpu_deg2_f64"] = round(res["deg2_f64_lu"]["stencils_per_s"] / rate, 1)
        log(f"weights CPU baseline: {res['cpu_baseline']}; GPU/CPU {res['gpu_over_cpu_deg2_f64']}x")
    return res


def stokes_case(ctx, R, level, n, budget, krylov, tol, cpu, cpu_iters):
    """One steady-Stokes solve size (see bench_stokes): FP64 and FP32-refinement GMRES over the
    same step budget, the operator and preconditioner applications, and optionally the CPU
    restatement's GMRES on its first cpu_iters steps."""
    xyz = R.nodes_icosahedral(level)
    xyz = np.ascontiguousarray(xyz[R.order_morton(xyz)])
    N = xyz.shape[0]
    st = R.knn_kdtree(xyz, n)
    off, idx = R.csr_from_fixed(st)
    eps = 0.035 * np.sqrt(N) - 0.1
    types = [R.LSFC, R.XSFC, R.YSFC, R.ZSFC]
    w, _ = ctx.weights(off, idx, xyz, eps, R.LSFC | R.XSFC | R.YSFC | R.ZSFC, dim=3, poly_deg=0)
    ws = [w[t].reshape(N, n) for t in range(4)]
    order = R.order_morton(xyz)
    e64 = R.EllOperators(ctx, st, ws, storage=R.F64, types=types)
    e32 = R.EllOperators(ctx, st, ws, storage=R.F32, types=types)
    for e in (e64, e32):
        e.set_row_order(order)
    ctx.sync()
    t0 = time.perf_counter()
    S = R.Stokes(ctx, e64, e32)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    rng = np.random.default_rng(0)

    def field():
        f = np.zeros(N)
        for _ in range(8):
            l = int(rng.integers(1, 11))
            f += rng.standard_normal() * R.sph_harm(l, int(rng.integers(-l, l + 1)), xyz)
        return f

    b = np.concatenate([field(), field(), field(), np.zeros(N), np.zeros(4)])
    d_b = ctx.to_device(b)
    S.gmres(d_b, tol, krylov, krylov)  # warm-up: code objects, workspace
    res = {}
    xs = {}
    for name, mode in (("fp64", R.F64), ("fp32_ir", R.F32)):
        ctx.sync()
        t = time.perf_counter()
        dx, it, err = S.gmres(d_b, tol, budget, krylov, mode, device_out=True)
        ctx.sync()
        sec = time.perf_counter() - t
        xs[name] = dx.download()
        dx.free()
        r = b - S.apply(xs[name])
        res[name] = {"s": round(sec, 3), "iterations": it, "ms_per_iteration": round(1e3 * sec / max(it, 1), 3),
                     "error": err, "converged": bool(err < tol),
                     "true_rel_residual": float(np.linalg.norm(r) / np.linalg.norm(b))}
        log(f"Stokes N={N} {name}: {it} GMRES steps in {sec:.2f}s, preconditioned residual {err:.2e}, "
            f"residual {res[name]['true_rel_residual']:.2e}")
    res["fp32_ir"]["rel_diff_from_fp64"] = float(np.linalg.norm(xs["fp32_ir"] - xs["fp64"]) / np.linalg.norm(xs["fp64"]))
    # operator and preconditioner applications (HIP events on the context stream)
    d_x = ctx.to_device(xs["fp64"])
    d_y = ctx.alloc(S.n, np.float64)
    times = {}
    for name, fn in (("apply_us", lambda: R.RBFX.rbfx_stokes_apply(S.h, d_x.ptr, d_y.ptr)),
                     ("precond_us", lambda: R.RBFX.rbfx_stokes_precond(S.h, d_y.ptr))):
        ctx.check(fn(), name)
        ctx.timer_start()
        for _ in range(20):
            ctx.check(fn(), name)
        times[name] = round(ctx.timer_stop() * 1e3 / 20, 1)
    d_x.free()
    d_y.free()
    d_b.free()
    out = {"N": N, "n": n, "unknowns": S.n, "eps": eps, "krylov": krylov, "tol": tol, "iteration_budget": budget,
           "sweep_steps": {"lower": S.lower_levels, "upper": S.upper_levels}, "setup_s": round(t_setup, 3),
           **res, **times}
    if cpu:
        O, build = cpu_oracle()
        A = O.stokes_assemble(st, np.stack(ws))
        F = O.ilu0(A, 0, 3 * N)
        t = time.perf_counter()
        _, _, it_c, err_c = O.gmres(A, b, F, tol=tol, max_iter=cpu_iters, krylov=krylov)
        sec_c = time.perf_counter() - t
        t = time.perf_counter()
        O.ilu0_apply(F, b)
        pre_c = time.perf_counter() - t
        out["cpu"] = {"ms_per_iteration": round(1e3 * sec_c / it_c, 2), "precond_ms": round(1e3 * pre_c, 2),
                      "cores": 1, "kind": "port",
                      "sample": f"the first {it_c} GMRES({krylov}) steps of the same system with the C++ restatement "
                                f"(oracle/oracle_stokes.cpp: ViennaCL's Householder GMRES, the reference's ILU(0); "
                                f"{build}, one thread)"}
        out["gpu_over_cpu_per_iteration"] = round(out["cpu"]["ms_per_iteration"] / res["fp64"]["ms_per_iteration"], 1)
    S.close()
    e64.close()
    e32.close()
    return out


def bench_stokes(ctx, R, cpu, small=True):
    """Steady Stokes on the sphere (the reference's StokesSteadyPDE with GMRES(50) + ILU(0) of the
    velocity block): icosahedral nodes, n = 31 kd-tree stencils, eps = 0.035 sqrt(N) - 0.1,
    eta = 1, a seeded body force of spherical harmonics of degree <= 10 on the velocity rows.
    Level 6 (N = 40962, the Stokes workload's size): a budget of 1000 steps (the reference's
    method stalls there), timed beside the CPU restatement. Level 3 (N = 642), where it
    converges: both precisions solved to 1e-8 and compared."""
    out = stokes_case(ctx, R, 6, 31, 1000, 50, 1e-8, cpu, 50)
    if small:
        out["converging_case"] = stokes_case(ctx, R, 3, 31, 4000, 50, 1e-8, False, 0)
    log(f"Stokes: {json.dumps(out)}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n-nodes", type=int, default=1000000)
    ap.add_argument("--width", type=int, default=50)
    ap.add_argument("--fields", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=1,
                    help="CPU baseline RK4 steps per timed run (median of 5 runs after one warm-up run)")
    ap.add_argument("--kernel-only", action="store_true",
                    help="set up, then launch only the stage kernels, the SpMV and the fused SpMM (profiling)")
    ap.add_argument("--no-weights", action="store_true", help="skip the batched uneven-stencil weights section")
    ap.add_argument("--no-revolution", action="store_true", help="skip the untimed one-revolution accuracy run")
    ap.add_argument("--weights-eps-scale", type=float, default=1.0,
                    help="with --weights-only: multiply the weights workload's eps (better conditioned when > 1)")
    ap.add_argument("--weights-only", action="store_true",
                    help="run only the batched uneven-stencil weights section (profiling)")
    ap.add_argument("--weights-f64-deg2", action="store_true",
                    help="with --weights-only: only the FP64 LU solve at degree 2 (one launch per size class)")
    ap.add_argument("--weights-sizes", default="17,31,50,75,101",
                    help="with --weights-only: comma-separated stencil sizes drawn uniformly (each of 17, 31, 50, 75, 101)")
    ap.add_argument("--no-stokes", action="store_true", help="skip the steady-Stokes solver section")
    ap.add_argument("--stokes-only", action="store_true", help="run only the steady-Stokes solver section")
    ap.add_argument("--stokes-no-small", action="store_true",
                    help="with --stokes-only: skip the converging N = 642 case (profiles of the N = 40962 case)")
    ap.add_argument("--node-order", action="store_true",
                    help="process rows in node order instead of the Morton locality order")
    ap.add_argument("--spiral-numbering", action="store_true",
                    help="keep the generator's (golden-angle spiral) node numbering instead of numbering the "
                         "nodes along the Morton curve")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the row-partitioned multi-GPU "
                                   "step is designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import rbffd_gpu_amd as R

    N, n, F = args.n_nodes, args.width, args.fields
    if args.stokes_only:
        bench_stokes(R.Context(0), R, not args.no_cpu, small=not args.stokes_no_small)
        return
    if args.weights_only:
        ctx = R.Context(0)
        sel = dict(degrees=(2,), modes=("f64_lu",)) if args.weights_f64_deg2 else {}
        size_set = tuple(int(k) for k in args.weights_sizes.split(","))
        if not size_set or any(k not in C3_EPS for k in size_set):
            sys.exit(f"--weights-sizes: each size must be one of {sorted(C3_EPS)}")
        res = bench_weights(ctx, R, N, False, 1, eps_scale=args.weights_eps_scale, size_set=size_set, **sel)
        log("weights-only done", json.dumps(res))
        return
    t0 = time.time()
    xyz = R.nodes_fibonacci(N, 1e-3, 0)
    xyz /= np.linalg.norm(xyz, axis=1)[:, None]
    if not args.spiral_numbering:
        # number the nodes along a Morton curve (the same point set, relabelled): the 64-byte field
        # records of a stencil's nodes then share cache lines and pages
        xyz = np.ascontiguousarray(xyz[R.order_morton(xyz)])
    t_nodes = time.time() - t0
    t0 = time.time()
    st = R.knn_kdtree(xyz, n)
    t_kdtree = time.time() - t0
    log(f"nodes {t_nodes:.1f}s, kd-tree stencils (host, all threads) {t_kdtree:.2f}s")
    ctx = R.Context(0)
    info = ctx.info()
    log("device", info)
    # the same stencils from the GPU search (rows with distance ties from the host kd-tree)
    R.knn_gpu(ctx, xyz[:1000].copy(), n)  # warm the kernel
    t0 = time.time()
    st_gpu, knn_host_rows = R.knn_gpu(ctx, xyz, n)
    t_knn_gpu = time.time() - t0
    # and straight into device memory (no copy of the lists back to the host)
    ctx.sync()
    t0 = time.time()
    d_st_gpu, _ = R.knn_gpu_device(ctx, xyz, n)
    ctx.sync()
    t_knn_dev = time.time() - t0
    knn = {"n_nodes": N, "k": n, "kdtree_host_s": round(t_kdtree, 3), "gpu_s": round(t_knn_gpu, 3),
           "gpu_to_device_s": round(t_knn_dev, 3), "host_rows": knn_host_rows,
           "identical_to_kdtree": bool(np.array_equal(st_gpu, st) and np.array_equal(d_st_gpu.download(), st)),
           "host_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))}
    d_st_gpu.free()
    del st_gpu
    log(f"GPU kNN stencils: {t_knn_gpu:.3f}s to host memory, {t_knn_dev:.3f}s to device memory "
        f"({knn_host_rows} rows from the host kd-tree), identical to the kd-tree: {knn['identical_to_kdtree']}")

    # ---- weights on the GPU ----------------------------------------------------------------
    mask = R.XSFC | R.YSFC | R.ZSFC | R.HV  # columns in ascending bit order: HV, XSFC, YSFC, ZSFC
    eps = 4.0 * (0.044 * np.sqrt(N) - 0.14)
    hv_k, hv_gamma = 4, 2.0 ** 9
    hv_scale = -hv_gamma * float(N) ** -hv_k
    d_xyz = ctx.to_device(xyz)
    off, idx = R.csr_from_fixed(st)
    d_off, d_idx = ctx.to_device(off), ctx.to_device(idx)
    ctx.sync()
    t0 = time.time()
    d_w, _ = ctx.weights(d_off, d_idx, d_xyz, eps, mask, dim=2, poly_deg=2 | R.POLY_SPHERE, hv_k=hv_k,
                         hv_scale=hv_scale, device_out=True)
    ctx.sync()
    t_weights = time.time() - t0
    log(f"GPU weights ({N} stencils, n={n}, 4 RHS): {t_weights:.3f}s")
    nnz = N * n
    d_st = d_idx.view(0, (N, n))
    w_views = [d_w.view(t * nnz, (N, n)) for t in range(4)]

    # ---- operators, fields, coefficients ------------------------------------------------------
    ell64 = R.EllOperators(ctx, d_st, w_views, storage=R.F64, types=R.dertype_bits(mask))
    ell32 = R.EllOperators(ctx, d_st, w_views, storage=R.F32, types=R.dertype_bits(mask))
    ell_sp64 = R.EllOperators(ctx, d_st, [w_views[1]], storage=R.F64)  # surface d/dx
    ell_sp32 = R.EllOperators(ctx, d_st, [w_views[1]], storage=R.F32)
    if not args.node_order:
        # rows processed along a Morton curve: a wave's rows share stencil nodes (results unchanged)
        t0 = time.time()
        order = R.order_morton(xyz)
        for e in (ell64, ell32):
            e.set_row_order(order)
        log(f"Morton row order: {time.time() - t0:.2f}s")
        t0 = time.time()
        # tiles and node caps: FP32 128-row tiles with 48 KB of staged records per block (three
        # 512-thread blocks per CU), FP64 128-row tiles with 78 KB (two per CU); the fastest of
        # sweeps over 32-, 64- and 128-row tiles and 16-78 KB caps
        for name, e, fb, rows, cap in (("fp64", ell64, 8, 128, 78 * 1024), ("fp32", ell32, 4, 128, 48 * 1024)):
            tl, k_max, n_tiles = e.build_local(rows, cap // (F * fb))
            log(f"LDS staging tables {name}: {n_tiles} tiles of <= {tl} rows, at most {k_max} distinct nodes per tile")
        log(f"LDS staging tables: {time.time() - t0:.2f}s")
    t0 = time.time()
    u0, spec = R.make_fields(xyz, F, seed=0)
    log(f"fields: {time.time() - t0:.1f}s")
    axis = np.array([0.3, -0.2, 1.0])
    omega = 2 * np.pi * axis / np.linalg.norm(axis)  # one revolution per unit time
    vel = np.cross(omega, xyz)
    coefs = [None, ctx.to_device(-vel[:, 0].copy()), ctx.to_device(-vel[:, 1].copy()), ctx.to_device(-vel[:, 2].copy())]
    dt = 1.0 / 4000
    d_u0 = ctx.to_device(u0)

    def timed(rk, K, W):
        rk.set_state(d_u0)
        rk.step(dt, W)
        ctx.sync()
        t = time.perf_counter()
        ctx.timer_start()
        rk.step(dt, K)
        ev_ms = ctx.timer_stop()
        wall = time.perf_counter() - t
        return wall, ev_ms

    rk32 = R.RK4(ctx, ell32, F, R.F32, R.F32, coefs=coefs)  # FP32 storage and accumulation, FP64 state
    rk64 = R.RK4(ctx, ell64, F, R.F64, R.F64, coefs=coefs)
    rk32a = R.RK4(ctx, ell32, F, R.F32, R.F64, coefs=coefs)  # FP64 accumulation of the gathers

    # ---- autotune the launch shapes (fields/thread, block, gather batch) ----------------------
    from rbffd_gpu_amd import tune as TU

    table = {}
    t0 = time.time()
    tuned = {}
    for name, rk, ell, ft, acc in (("fp32", rk32, ell32, R.F32, R.F32), ("fp64", rk64, ell64, R.F64, R.F64),
                                   ("fp32_acc64", rk32a, ell32, R.F32, R.F64)):
        rk.set_state(d_u0)
        entry, res = TU.tune_rk4(ctx, rk, ell, ft, acc, dt=dt, table=table)
        tuned[name] = entry
        log(f"tuned {name} RK4 stage shape: {entry}; worst {res[-1][0] * 1e3:.0f} us")
        log("  fastest shapes (us/step, fpt, block, batch, share): " +
            "; ".join(f"{r[0] * 1e3:.0f} {r[1:]}" for r in res[:6]))
    for name, ell, ft, dt_np in (("spmv_fp64", ell_sp64, R.F64, np.float64), ("spmv_fp32", ell_sp32, R.F32, np.float32)):
        d_x1 = ctx.to_device(u0[:, 0].astype(dt_np))
        d_y1 = ctx.alloc((1, N, 1), dt_np)
        entry, _ = TU.tune_spmm(ctx, ell, 1, d_x1, d_y1, ft, R.F64, table=table)
        tuned[name] = entry
        d_x1.free()
        d_y1.free()
    log(f"autotune: {time.time() - t0:.1f}s {tuned}")

    if args.kernel_only:
        rk32.set_state(d_u0)
        rk64.set_state(d_u0)
        for _ in range(20):
            for s in (1, 2, 3, 4):
                rk32.stage(s, dt)
        for _ in range(20):
            for s in (1, 2, 3, 4):
                rk64.stage(s, dt)
        rk32a.set_state(d_u0)
        for _ in range(20):
            for s in (1, 2, 3, 4):
                rk32a.stage(s, dt)
        d_x = ctx.to_device(u0[:, 0].copy())
        d_y = ctx.alloc(N, np.float64)
        for _ in range(20):
            ell_sp64.spmv(0, d_x, d_y)
        # the fused SpMM (4 operators x F fields, every output written), tuned, then 20 launches
        for ell, dt_np, ft in ((ell64, np.float64, R.F64), (ell32, np.float32, R.F32)):
            d_xs = ctx.to_device(u0.astype(dt_np))
            d_ys = ctx.alloc((4, N, F), dt_np)
            TU.tune_spmm(ctx, ell, F, d_xs, d_ys, ft, ft, table=table)
            for _ in range(20):
                ell.spmm(F, d_xs, d_ys, ft, ft)
            ctx.sync()
            d_xs.free()
            d_ys.free()
        stream_copy_gbs(ctx, R, reps=20)
        ctx.sync()
        log("kernel-only launches done")
        return

    K, W = args.steps, args.warmup
    wall32, ev32 = timed(rk32, K, W)
    u32 = rk32.get_state()
    wall64, ev64 = timed(rk64, K, W)
    u64 = rk64.get_state()
    wall32a, _ = timed(rk32a, K, W)
    u32a = rk32a.get_state()
    log(f"FP32-stored: {K / wall32:.1f} steps/s (wall) {K / (ev32 / 1e3):.1f} (events); "
        f"FP64: {K / wall64:.1f} steps/s (wall) {K / (ev64 / 1e3):.1f} (events)")

    # deviation FP32 vs FP64 and error vs the analytic rotated field after W+K steps
    dev = float(np.abs(u32 - u64).max() / np.abs(u64).max())
    dev_rms = float(np.sqrt(np.mean((u32 - u64) ** 2) / np.mean(u64 ** 2)))
    t_end = (W + K) * dt
    xr = rotate(xyz, omega, -np.linalg.norm(omega) * t_end)
    exact = np.zeros_like(u0)
    for f in range(F):
        for l, m, c in spec[f]:
            exact[:, f] += c * R.sph_harm(int(l), int(m), xr)
    err64 = float(np.abs(u64 - exact).max() / np.abs(exact).max())
    err32 = float(np.abs(u32 - exact).max() / np.abs(exact).max())
    dev32a = float(np.abs(u32a - u64).max() / np.abs(u64).max())
    dev32a_rms = float(np.sqrt(np.mean((u32a - u64) ** 2) / np.mean(u64 ** 2)))
    log(f"FP32-stored with FP64 accumulation: {K / wall32a:.1f} steps/s, deviation {dev32a:.2e} (max) {dev32a_rms:.2e} (rms)")
    log(f"rel. deviation FP32 vs FP64 {dev:.2e} (max), {dev_rms:.2e} (rms); max rel error vs analytic: FP64 {err64:.2e}, FP32 {err32:.2e}")

    # accuracy over one full revolution (untimed): FP32 vs FP64 and both vs the analytic field,
    # which is then the initial field again
    rev = {}
    if not args.no_revolution:
        n_rev = int(round(1.0 / dt))
        ends = {}
        for name, rk in (("fp32", rk32), ("fp32_acc64", rk32a), ("fp64", rk64)):
            rk.set_state(d_u0)
            rk.step(dt, n_rev)
            ends[name] = rk.get_state()
        r32, r32a, r64 = ends["fp32"], ends["fp32_acc64"], ends["fp64"]
        rev = {"steps": n_rev,
               "fp32_rel_deviation_from_fp64": float(np.abs(r32 - r64).max() / np.abs(r64).max()),
               "fp32_rms_rel_deviation_from_fp64": float(np.sqrt(np.mean((r32 - r64) ** 2) / np.mean(r64 ** 2))),
               "fp32_acc64_rel_deviation_from_fp64": float(np.abs(r32a - r64).max() / np.abs(r64).max()),
               "fp32_acc64_rms_rel_deviation_from_fp64": float(np.sqrt(np.mean((r32a - r64) ** 2) /
                                                                       np.mean(r64 ** 2))),
               "max_rel_error_vs_analytic": {k: float(np.abs(v - u0).max() / np.abs(u0).max()) for k, v in ends.items()},
               "l2_rel_error_vs_analytic": {k: float(np.sqrt(np.mean((v - u0) ** 2) / np.mean(u0 ** 2)))
                                            for k, v in ends.items()}}
        log(f"one revolution ({n_rev} steps): deviation FP32 vs FP64 {rev['fp32_rel_deviation_from_fp64']:.2e} (max) "
            f"{rev['fp32_rms_rel_deviation_from_fp64']:.2e} (rms), with FP64 accumulation "
            f"{rev['fp32_acc64_rel_deviation_from_fp64']:.2e} (max); error vs analytic max {rev['max_rel_error_vs_analytic']}, "
            f"l2 {rev['l2_rel_error_vs_analytic']}")

    # ---- dominant kernel: the fused stage kernel, timed per launch ---------------------------
    R_L = 50
    stage_ms = {}
    for name, rk, s_w, s_f in (("fp32", rk32, 4, 4), ("fp64", rk64, 8, 8)):
        rk.set_state(d_u0)
        for s in (1, 2, 3, 4):
            rk.stage(s, dt)
        ctx.sync()
        per = []
        for s in (1, 2, 3, 4):
            ctx.timer_start()
            for _ in range(R_L):
                rk.stage(s, dt)
            per.append(ctx.timer_stop() / R_L)
        stage_ms[name] = per
    bytes32 = [bytes_per_stage(N, n, 4, F, 4, 4, 3, s) for s in (1, 2, 3, 4)]
    bytes64 = [bytes_per_stage(N, n, 4, F, 8, 8, 3, s) for s in (1, 2, 3, 4)]
    avg_ms32 = float(np.mean(stage_ms["fp32"]))
    avg_b32 = float(np.mean(bytes32))
    ach32 = avg_b32 / (avg_ms32 * 1e-3) / 1e9
    ach64 = float(np.mean(bytes64)) / (float(np.mean(stage_ms["fp64"])) * 1e-3) / 1e9
    t32 = tuned["fp32"]
    sg32 = F // t32["fpt"] if t32.get("share") else 0
    pd32 = t32["share"] if t32.get("share", 0) >= 2 else 0
    if t32.get("share") in (4, 5):  # LDS-staged kernel (prefetch depth 2; 5: matrix-core multiply-adds)
        mx = "true" if t32["share"] == 5 else "false"
        kname32 = f"rk4_stage_lds_kernel<float, float, float, 4, {t32['fpt']}, {t32['batch']}, {sg32}, 2, {mx}>"
    else:
        kname32 = f"rk4_stage_kernel<float, float, float, 4, {t32['fpt']}, {t32['batch']}, {sg32}, {pd32}>"
    traffic32, traffic_kernel = profile_traffic(kname32)
    log(f"stage kernel avg: fp32 {avg_ms32 * 1e3:.1f} us ({ach32:.0f} GB/s), "
        f"fp64 {np.mean(stage_ms['fp64']) * 1e3:.1f} us ({ach64:.0f} GB/s)")

    # ---- SpMV (1 operator, 1 field) ------------------------------------------------------------
    spmv = {}
    for name, ell, dt_np, ft, s_w in (("fp64", ell_sp64, np.float64, R.F64, 8), ("fp32", ell_sp32, np.float32, R.F32, 4)):
        d_x = ctx.to_device(u0[:, 0].astype(dt_np))
        d_y = ctx.alloc(N, dt_np)
        for _ in range(5):
            ell.spmv(0, d_x, d_y, ft, R.F64)
        ctx.timer_start()
        for _ in range(R_L):
            ell.spmv(0, d_x, d_y, ft, R.F64)
        ms = ctx.timer_stop() / R_L
        b = N * n * (s_w + 4) + N * 2 * np.dtype(dt_np).itemsize
        spmv[name] = {"us": ms * 1e3, "GB_s": b / (ms * 1e-3) / 1e9, "bytes": b}
    log(f"SpMV: {spmv}")

    # ---- fused SpMM (4 operators x F fields, standalone, every output written) ------------------
    # algorithmic bytes: ops N n s_w + 4 N n (indices) + F N s_f (gathered fields, once) + ops F N s_out
    spmm = {}
    for name, ell, dt_np, ft, s_w in (("fp64", ell64, np.float64, R.F64, 8), ("fp32", ell32, np.float32, R.F32, 4)):
        d_x = ctx.to_device(u0.astype(dt_np))
        d_y = ctx.alloc((4, N, F), dt_np)
        entry, _ = TU.tune_spmm(ctx, ell, F, d_x, d_y, ft, ft, table=table)
        for _ in range(3):
            ell.spmm(F, d_x, d_y, ft, ft)
        ctx.timer_start()
        for _ in range(R_L):
            ell.spmm(F, d_x, d_y, ft, ft)
        ms = ctx.timer_stop() / R_L
        s_f = np.dtype(dt_np).itemsize
        b = 4 * N * n * s_w + 4 * N * n + F * N * s_f + 4 * F * N * s_f
        spmm[name] = {"us": round(ms * 1e3, 2), "GB_s": round(b / (ms * 1e-3) / 1e9, 1), "bytes": int(b),
                      "frac_of_hbm_peak": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "flop_per_launch": int(2 * 4 * F * N * n), "tuned": entry}
        d_x.free()
        d_y.free()
    log(f"SpMM (4 ops x {F} fields): {spmm}")
    copy_gbs = stream_copy_gbs(ctx, R)
    log(f"plain copy kernel: {copy_gbs:.0f} GB/s")

    # ---- batched weights with uneven stencils (own workload, reported in extra) ----------------
    host_threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    weights_c3 = None
    if not args.no_weights:
        for a in (ell64, ell32, ell_sp64, ell_sp32):  # the RK4 measurements are done: free HBM
            a.close()
        weights_c3 = bench_weights(ctx, R, N, not args.no_cpu, host_threads)
    stokes = None if args.no_stokes else bench_stokes(ctx, R, not args.no_cpu)

    # ---- CPU baseline: the oracle's RK4 step on the same workload (all host cores) --------------
    cpu = None
    if not args.no_cpu:
        O, build = cpu_oracle()
        # the host's thread budget: OMP_NUM_THREADS when the environment sets it, else the affinity mask
        nthreads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        w_host = [v.download() for v in w_views]
        coefs_h = [None, -vel[:, 0].copy(), -vel[:, 1].copy(), -vel[:, 2].copy()]
        steps_per_run = max(1, args.cpu_steps)

        stp = O.RK4Stepper(st, w_host, coefs_h, F, n_threads=nthreads)  # weights stacked, workspace allocated once
        u_cpu = np.array(u0, np.float64, order="C")

        def leg(threads, k):
            def f():
                stp.n_threads = threads
                t = time.perf_counter()
                stp.step(u_cpu, dt, k)
                return k, time.perf_counter() - t
            return f

        rate, rates = median_rate(leg(nthreads, steps_per_run), runs=5, warmup=1)
        cpu = {"value": round(rate, 4), "unit": "RK4 steps/s", "cores": nthreads, "kind": "port",
               "sample": f"FP64 RK4 steps of the same N={N}, n={n}, 4-operator, {F}-field workload with the C++ "
                         f"restatement (oracle/oracle.cpp, {build}), OpenMP over rows, workspace allocated once; "
                         f"median of 5 timed runs of {steps_per_run} step(s) after one warm-up run",
               "runs_steps_per_s": [round(r, 4) for r in rates],
               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
        r1, _ = median_rate(leg(1, 1), runs=3, warmup=0)
        cpu["one_thread"] = {"value": round(r1, 4), "sample": "1 FP64 RK4 step on one thread, median of 3 runs"}
        log(f"CPU baseline: {cpu}")

    value = K / wall32
    out = {
        "metric": "RK4 steps/s + SpMV HBM GB/s at N=1M,n=50; FP32 vs FP64 at stated error",
        "value": round(value, 2),
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * wall32 / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage and accumulation / f64 state",
        "data": "synthetic (seeded sphere nodes, spherical-harmonic fields)",
        "config": {"workload": f"fused RK4 step: N={N}, n={n} kd-tree stencils, 4 operators (surface d/dx,d/dy,d/dz + HV Delta^4) x "
                               f"{F} fields, FP32-stored weights+fields, FP32 accumulation, FP64 state, hipGraph",
                   "N": N, "n": n, "operators": 4, "fields": F, "dt": dt, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach32 / HBM_PEAK_GBS, 4), "traffic": traffic32,
                     "kernel": kname32, "algorithmic_bytes_per_launch": int(avg_b32),
                     "avg_launch_us": round(avg_ms32 * 1e3, 2),
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/rk4_stage_N1M_n50_ops4_F16.md",
                     "traffic_kernel": traffic_kernel},
        "cpu_baseline": cpu,
        "extra": {
            "fp64_steps_per_s": round(K / wall64, 2),
            "fp32_acc64": {"steps_per_s": round(K / wall32a, 2), "rel_deviation_from_fp64": dev32a,
                           "rms_rel_deviation_from_fp64": dev32a_rms},
            "fp32_over_fp64": round(wall64 / wall32, 3),
            "fp32_rel_deviation_from_fp64": dev,
            "fp32_rms_rel_deviation_from_fp64": dev_rms,
            "rel_error_vs_analytic": {"fp64": err64, "fp32": err32, "t": t_end},
            "one_revolution": rev,
            "step_roofline_frac": {"fp32": round(sum(bytes32) / (wall32 / K) / 1e9 / HBM_PEAK_GBS, 4),
                                   "fp64": round(sum(bytes64) / (wall64 / K) / 1e9 / HBM_PEAK_GBS, 4)},
            "stage_us": {k: [round(x * 1e3, 2) for x in v] for k, v in stage_ms.items()},
            "stage_GB_s_fp64": round(ach64, 1),
            "copy_kernel_GB_s": round(copy_gbs, 1),
            "spmv": spmv,
            "spmm_4ops": spmm,
            "weights_s": round(t_weights, 4),
            "weights_uneven": weights_c3,
            "stokes": stokes,
            "knn": knn,
            "tuned_shapes": tuned,
            "device": info,
        },
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
