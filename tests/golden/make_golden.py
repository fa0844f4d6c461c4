"""Writes the committed golden fixtures under tests/golden/. Run in the build container,
where the reference tree (for its kd-tree) and scipy (LAPACK) are available:

    make -C oracle && python tests/golden/make_golden.py

Fixtures (data only):
  c1_knn.npz       config-1 nodes (N=4096 Fibonacci +-1e-3, seed 0) and the reference
                   kd-tree's n=31 stencils (oracle/_ref/ref_knn, compiled from
                   src/grids/grid_interface.cpp + external/kdtree/KDTree.h)
  ties_knn.npz     tie-heavy node sets: the reference's own 10x10 grid fixture
                   (tests/unified/regulargrid_10x_10y_1z_final.ascii), a 7x6x5 integer
                   lattice, and the grid with 10 duplicated nodes; reference stencils for
                   several k
  large_knn.json   sha256 digests of the node sets and reference stencils at N=163842
                   (n=101) and N=1e6 (n=50), plus the first rows of each
  lapack_weights.npz  stencil systems solved by scipy.linalg.solve (LAPACK getrf/getrs,
                   the routine arma::solve reaches): node coordinates, eps, mask, degree,
                   the solution and cond(M)
  regulargrid_10x_10y_1z_final.ascii   the reference's own node fixture (a data file of
                   its tests/unified/), copied unchanged
  stencils_maxsz<k>_regulargrid_10x_10y_1z_final.ascii   the stencil files the reference's
                   Grid::writeStencilsToFile writes for that grid (k = 9, 13), produced by
                   oracle/_ref/ref_knn from the compiled reference sources

`python tests/golden/make_golden.py stencil_files` rewrites only the last two.
"""
import hashlib
import json
import os
import shutil
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
import rbffd_gpu_amd as R  # noqa: E402

REF_GRID = "/root/reference/tests/unified/regulargrid_10x_10y_1z_final.ascii"


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def stencil_files():
    """The reference node fixture and the reference's own stencil files for it."""
    name = os.path.basename(REF_GRID)
    shutil.copyfile(REF_GRID, os.path.join(HERE, name))
    grid = np.loadtxt(REF_GRID)[:, :3]
    with tempfile.TemporaryDirectory() as td:
        fn = os.path.join(td, "n.bin")
        with open(fn, "wb") as f:
            f.write(struct.pack("<q", grid.shape[0]))
            f.write(grid.tobytes())
        for k in (9, 13):
            subprocess.run([O.REF_KNN, fn, str(k), os.path.join(td, "o.bin"), name], cwd=td, check=True,
                           stdout=subprocess.DEVNULL)
            out = f"stencils_maxsz{k}_{name}"
            shutil.copyfile(os.path.join(td, out), os.path.join(HERE, out))


def main():
    import scipy.linalg as sl

    stencil_files()

    # ---- config 1 -----------------------------------------------------------------------
    xyz = R.nodes_fibonacci(4096, 1e-3, 0)
    st = O.ref_knn(xyz, 31)
    np.savez_compressed(os.path.join(HERE, "c1_knn.npz"), xyz=xyz, st_ref=st)

    # ---- tie-heavy sets -----------------------------------------------------------------
    grid = np.loadtxt(REF_GRID)[:, :3]
    lat = np.array([[i, j, k] for i in range(7) for j in range(6) for k in range(5)], float)
    dup = np.vstack([grid, grid[:10]])
    out = {"grid": grid, "lattice": lat, "dups": dup}
    for name, pts, ks in (("grid", grid, (5, 9, 13, 31)), ("lattice", lat, (7, 27, 50)), ("dups", dup, (9, 17))):
        for k in ks:
            out[f"{name}_k{k}"] = O.ref_knn(pts, k)
    np.savez_compressed(os.path.join(HERE, "ties_knn.npz"), **out)

    # ---- large sets: digests -------------------------------------------------------------
    big = {}
    for n, k, seed in ((163842, 101, 0), (1000000, 50, 0)):
        x = R.nodes_fibonacci(n, 1e-3, seed)
        s = O.ref_knn(x, k)
        big[f"N{n}_k{k}"] = {"n": n, "k": k, "seed": seed, "jitter": 1e-3, "xyz_sha256": sha(x),
                             "st_sha256": sha(s.astype(np.int32)), "first_rows": s[:16].tolist()}
    with open(os.path.join(HERE, "large_knn.json"), "w") as f:
        json.dump(big, f, indent=1)

    # ---- LAPACK-solved systems ----------------------------------------------------------
    rng = np.random.default_rng(1)
    xs = R.nodes_fibonacci(20000, 1e-3, 1)
    eps_tab = {17: 0.026 * np.sqrt(20000) - 0.08, 31: 0.035 * np.sqrt(20000) - 0.1, 50: 0.044 * np.sqrt(20000) - 0.14,
               75: 0.051 * np.sqrt(20000) - 0.15, 101: 0.058 * np.sqrt(20000) - 0.16}
    systems = []
    c1_ids = rng.choice(4096, 24, replace=False)
    for i in c1_ids:
        systems.append((xyz[st[i]], 3.5, R.X | R.Y | R.Z | R.LSFC, 3, 2, 0))
    stc = {n: R.knn_kdtree(xs, n) for n in eps_tab}
    for deg in (0, 2, 3):
        for n in eps_tab:
            for i in rng.choice(20000, 3, replace=False):
                systems.append((xs[stc[n][i]], eps_tab[n], R.X | R.Y | R.Z | R.LAPL, 3, deg, 0))
    # cosine-bell operators (LAMBDA, THETA, HV dim 2), the reference's constant monomial
    st101 = stc[101]
    for i in rng.choice(20000, 8, replace=False):
        systems.append((xs[st101[i]], 0.035 * np.sqrt(20000), R.LAMBDA | R.THETA | R.HV, 2, 0, 4))
    flat = {}
    for q, (pts, eps, mask, dim, deg, hvk) in enumerate(systems):
        n = pts.shape[0]
        deg = O.effective_degree(deg, n)
        M, Rr = O.stencil_system(pts, np.arange(n, dtype=np.int32), eps, mask, dim, deg, hvk)
        Xs = sl.solve(M, Rr)
        flat[f"s{q}_xyz"] = pts
        flat[f"s{q}_meta"] = np.array([eps, mask, dim, deg, hvk, np.linalg.cond(M)])
        flat[f"s{q}_x"] = Xs
    flat["count"] = np.array(len(systems))
    np.savez_compressed(os.path.join(HERE, "lapack_weights.npz"), **flat)
    print("wrote fixtures:", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    if sys.argv[1:] == ["stencil_files"]:
        stencil_files()
    else:
        main()
