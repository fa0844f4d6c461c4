"""Node and stencil files through the C++ drop-in layer (GridReader, Grid stencil files,
Grid::computeStencilRadii), on the CPU, against the reference's own outputs:

* the node fixture is the reference's tests/unified/regulargrid_10x_10y_1z_final.ascii;
* the golden stencil files were written by the reference's Grid::writeStencilsToFile
  (grid_interface.cpp:144-170), compiled from its sources (tests/golden/make_golden.py).
"""
import json
import os
import shutil
import subprocess

import numpy as np

from conftest import GOLDEN, ROOT

BIN = os.path.join(ROOT, "rbffd_gpu_amd", "bin", "rbfx_stencil_generator")
GRID = "regulargrid_10x_10y_1z_final.ascii"


def run(args, cwd, ok=True):
    p = subprocess.run([BIN] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=120)
    if ok:
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        return json.loads(p.stdout.strip().splitlines()[-1])
    return p


def radii_ref(xyz, st):
    d = np.sqrt(((xyz[st] - xyz[st[:, :1]]) ** 2).sum(2))
    return d[:, 1:].mean(1), d[:, 1], d[:, -1]


def test_gridreader_and_written_stencil_file_match_reference(tmp_path):
    shutil.copy(os.path.join(GOLDEN, GRID), tmp_path)
    xyz = np.loadtxt(os.path.join(GOLDEN, GRID))
    for k in (9, 13):
        res = run([GRID, 3, 100, k], tmp_path)
        assert res["generated"] and res["nodes"] == 100 and res["stencils"] == 100 and res["max_size"] == k
        fname = f"stencils_maxsz{k}_{GRID}"
        assert res["file"] == fname
        # byte-identical to the file the reference writes for the same grid (ties included)
        assert (tmp_path / fname).read_bytes() == open(os.path.join(GOLDEN, fname), "rb").read()
        st = np.array([[int(v) for v in line.split()[1:]] for line in open(os.path.join(GOLDEN, fname))])
        avg, mn, mx = np.fromfile(tmp_path / f"radii_{GRID}.bin").reshape(3, 100)
        ra, rn, rx = radii_ref(xyz, st)
        np.testing.assert_allclose(avg, ra, rtol=1e-15)
        assert np.array_equal(mn, rn) and np.array_equal(mx, rx)


def test_reference_stencil_file_is_loaded(tmp_path):
    """A stencil file in the reference's layout is read back as the reference wrote it (the
    leading size is not taken for an index), and its radii are recomputed."""
    shutil.copy(os.path.join(GOLDEN, GRID), tmp_path)
    fname = f"stencils_maxsz13_{GRID}"
    # mark the file so that a regenerated one would differ: swap two ranks of stencil 0
    lines = open(os.path.join(GOLDEN, fname)).read().splitlines()
    row0 = lines[0].split()
    row0[2], row0[3] = row0[3], row0[2]
    lines[0] = " ".join(row0)
    (tmp_path / fname).write_text("\n".join(lines) + "\n")
    res = run([GRID, 3, 100, 13], tmp_path)
    assert not res["generated"] and res["stencils"] == 100
    st = np.array([[int(v) for v in line.split()[1:]] for line in lines])
    avg, mn, mx = np.fromfile(tmp_path / f"radii_{GRID}.bin").reshape(3, 100)
    xyz = np.loadtxt(os.path.join(GOLDEN, GRID))
    ra, rn, rx = radii_ref(xyz, st)
    np.testing.assert_allclose(avg, ra, rtol=1e-15)
    assert np.array_equal(mn, rn) and np.array_equal(mx, rx)


def test_gridreader_node_count_and_extra_columns(tmp_path):
    xyz = np.loadtxt(os.path.join(GOLDEN, GRID))
    # a 4-column file ("x y z extra"): the extra column is skipped
    four = np.column_stack([xyz, np.arange(100.0)])
    np.savetxt(tmp_path / "four.ascii", four, fmt="%.17g")
    res = run(["four.ascii", 4, 100, 9], tmp_path)
    assert res["nodes"] == 100
    got = [[int(v) for v in line.split()[1:]] for line in open(tmp_path / "stencils_maxsz9_four.ascii")]
    ref = [[int(v) for v in line.split()[1:]] for line in open(os.path.join(GOLDEN, f"stencils_maxsz9_{GRID}"))]
    assert got == ref
    # at most grid_size nodes are read
    shutil.copy(os.path.join(GOLDEN, GRID), tmp_path)
    res = run([GRID, 3, 40, 5], tmp_path)
    assert res["nodes"] == 40
    # asking for more nodes than the file holds is an error (print + exit), as in the reference
    p = run([GRID, 3, 200, 5], tmp_path, ok=False)
    assert p.returncode != 0 and "Found only 100 nodes" in p.stdout
