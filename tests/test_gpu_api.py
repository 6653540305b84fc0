"""The drop-in surface around the hot path on the GPU:

* per-cell refinement counts (NumericalIntegrator3D::getRefinementsRequired)
  against the oracle's restatement of the reference's call sequence, both
  integration paths;
* the fixed-level refined mesh (getRefinedVertices / getRefinedCells /
  getRefinedCellMeasures) against what the midpoint split fixes: exact
  midpoints, the reference's child order, quarter measures, children that
  tile their parent with its orientation;
* the device task lists (Evaluator3D::getTasks, Mesh3D::getNotNeighbors);
* the command-line driver on the reference's class API, end to end: results
  files identical to the library's, VTK refinement counts equal to the API's,
  OBJ/VTK files of the refined mesh.
"""
import ctypes
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from tests.helpers import runge_criterion
from tests.test_gpu_fixtures import _write_dat
from tests.test_gpu_parity import run_gpu

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "integrator2_amd", "integrator2test3D")


def mesh2(i2):
    return i2.icosphere(2, (2.0, 1.0, 0.7), 0.05, 0)


@pytest.mark.parametrize("path", [0, 1])
def test_refinement_counts_match_oracle(gpu, oracle_mod, path):
    xyz, tri = mesh2(gpu)
    maxL, tol = 2, 1e-8
    (m, it, ev), out, Ks = run_gpu(gpu, xyz, tri, maxLevel=maxL, tol=tol, path=path, check=False)
    orc = oracle_mod.Oracle(xyz, tri)
    for t in range(3):
        got = it.getRefinementsRequired(t)
        tasks = out[t]["tasks"]
        print(f"path {path} type {t}: {len(tasks)} tasks, counts {np.bincount(got)}", flush=True)
        I4, J3, conv, K, lev, unconv = orc.run_tasks(t, tasks, gpu.QF3D13_L12, gpu.QF3D13_W, 7, mode=1, level=maxL,
                                                      tol=tol, want_levels=True, want_unconv=True)
        ref = orc.refinement_counts(tasks, conv, unconv, maxL)
        assert got.shape == ref.shape
        bad = np.nonzero(got != ref)[0]
        # a cell may differ only through a task whose Runge criterion lies within
        # its rounding noise of the tolerance (a decision the parity tests allow)
        for i in bad:
            sel = np.nonzero(tasks[:, 0] == i)[0]
            crit = np.stack([runge_criterion(lev[sel, L], lev[sel, L - 1]) for L in range(1, maxL + 1)], 1)
            assert (np.abs(crit - tol) / tol <= 1e-2).any(), (t, i, got[i], ref[i])
        assert len(bad) <= 2, (t, bad)
        has_task = np.bincount(tasks[:, 0], minlength=len(tri)) > 0
        assert ((got >= 1) == has_task).all() and got.max() <= maxL + 1


def test_refinement_counts_level0_and_fixed(gpu):
    xyz, tri = mesh2(gpu)
    (m, it, ev), out, _ = run_gpu(gpu, xyz, tri, maxLevel=0, tol=1e-8, path=1, check=False)
    for t in range(3):
        has_task = np.bincount(out[t]["tasks"][:, 0], minlength=len(tri)) > 0
        assert np.array_equal(it.getRefinementsRequired(t), has_task.astype(np.uint8))
    it.setFixedRefinementLevel(1)
    ev.runAllPairs(False)
    assert all(len(it.getRefinementsRequired(t)) == 0 for t in range(3))


def tri_area(xyz, tri):
    a, b, c = xyz[tri[:, 0]], xyz[tri[:, 1]], xyz[tri[:, 2]]
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1), np.cross(b - a, c - a)


def test_refined_mesh_fixed_level(gpu):
    xyz, tri = gpu.icosphere(1, (2.0, 1.0, 0.7), 0.05, 1)
    V, T = len(xyz), len(tri)
    m = gpu.Mesh3D()
    m.setData(xyz, tri)
    m.prepareMesh()
    it = gpu.NumericalIntegrator3D(m)
    with pytest.raises(gpu.I2Error):
        it.refinedMesh()                       # automatic mode keeps no refined mesh
    area0, n0 = tri_area(xyz, tri)
    assert np.allclose(area0, m.getCellMeasures(), rtol=1e-14, atol=0)
    # level 1: the reference's slots and child order
    it.setFixedRefinementLevel(1)
    rv, rc, rs = it.refinedMesh()
    assert rv.shape == (V + 3 * T, 3) and rc.shape == (4 * T, 3)
    assert np.array_equal(rv[:V], xyz)
    A, B, C = xyz[tri[:, 0]], xyz[tri[:, 1]], xyz[tri[:, 2]]
    assert np.array_equal(rv[V::3], 0.5 * (B + C))
    assert np.array_equal(rv[V + 1::3], 0.5 * (C + A))
    assert np.array_equal(rv[V + 2::3], 0.5 * (A + B))
    bc, ca, ab = V + 3 * np.arange(T), V + 3 * np.arange(T) + 1, V + 3 * np.arange(T) + 2
    expect = np.stack([np.stack([ab, tri[:, 1], bc], 1), np.stack([bc, tri[:, 2], ca], 1),
                       np.stack([ca, tri[:, 0], ab], 1), np.stack([bc, ca, ab], 1)], 1).reshape(-1, 3)
    assert np.array_equal(rc, expect)
    assert np.array_equal(rs, np.repeat(0.25 * m.getCellMeasures(), 4))
    # level 3: 64 descendants per cell tile the parent with its orientation
    it.setFixedRefinementLevel(3)
    rv, rc, rs = it.refinedMesh()
    assert rv.shape == (V + 63 * T, 3) and rc.shape == (64 * T, 3)
    assert np.array_equal(rs, np.repeat(m.getCellMeasures() / 64.0, 64))
    a3, n3 = tri_area(rv, rc)
    assert np.allclose(a3.reshape(T, 64).sum(1), area0, rtol=1e-13, atol=0)
    assert (np.einsum("cki,ci->ck", n3.reshape(T, 64, 3), n0) > 0).all()
    # every descendant vertex has barycentric coordinates k/8 in its parent
    P = rv[rc].reshape(T, 64 * 3, 3)
    M = np.stack([B - A, C - A], 2)                      # (T, 3, 2)
    for c in range(T):
        lam = np.linalg.lstsq(M[c], (P[c] - A[c]).T, rcond=None)[0]
        assert np.abs(lam * 8 - np.round(lam * 8)).max() < 1e-9
        assert (np.round(lam * 8) >= 0).all() and (np.round(lam * 8).sum(0) <= 8).all()


def test_device_task_lists(gpu):
    xyz, tri = mesh2(gpu)
    (m, it, ev), out, _ = run_gpu(gpu, xyz, tri, fixed=0, check=False)
    for t in range(3):
        assert np.array_equal(ev.getTasks(t), out[t]["tasks"]), t
    d = ctypes.c_void_p()
    n = ctypes.c_int64()
    gpu._check(gpu.lib().i2_mesh_device_array(m._h, ctypes.c_int(6), ctypes.byref(d), ctypes.byref(n)))
    assert np.array_equal(gpu._device_to_host(d, (n.value, 3), np.int32), m.getNotNeighbors())


def run_driver(args, cwd):
    print("driver:", " ".join(args), flush=True)
    r = subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def read_csv(path):
    lines = open(path).read().split("\n")
    assert lines[-1] == ""
    return lines[0], [l.split(";") for l in lines[1:-1]]


def vtk_cell_arrays(path):
    root = ET.fromstring(open(path).read().split("\n", 1)[1])
    return {a.get("Name"): np.array(a.text.split(), int) for a in root.findall("PolyData/Piece/CellData/DataArray")}


def test_driver_end_to_end(gpu, tmp_path):
    assert os.path.exists(DRIVER)
    xyz, tri = mesh2(gpu)
    dat = tmp_path / "ellipsoid.dat"
    _write_dat(str(dat), xyz, tri)
    run_auto = tmp_path / "auto"
    run_auto.mkdir()
    out = run_driver(["-f", str(dat), "-c", "--exporttocsv", "--exporttovtk", "--maxlevel=2", "--tolerance=1e-8"], str(run_auto))
    assert "Using adaptive error control procedure" in out and "Loaded mesh with 162 vertices and 320 cells" in out
    # the same run through the library's Python binding
    m = gpu.Mesh3D()
    assert m.loadMeshFromFile(str(dat))
    m.prepareMesh()
    it = gpu.NumericalIntegrator3D(m)
    it.setErrorControl(2, 1e-8)
    ev = gpu.EvaluatorJ3DK(m, it)
    ev.runAllPairs(True)
    names = ["SimpleNeighbors", "AttachedNeighbors", "NotNeighbors"]
    for t in range(3):
        r = ev.results(t, errors=True)
        header, rows = read_csv(run_auto / f"{names[t]}.csv")
        assert header == '"TaskI";"TaskJ";"IntegralX";"IntegralY";"IntegralZ";"Error"'
        assert len(rows) == len(r["tasks"])
        expect = [[str(a), str(b), format(x, "g"), format(y, "g"), format(z, "g"), format(e, "g")]
                  for (a, b, _), (x, y, z), e in zip(r["tasks"], r["results"], r["errors"])]
        assert rows == expect, t
    arrays = vtk_cell_arrays(run_auto / "OriginalMesh.vtp")
    assert sorted(arrays) == sorted(f"{n}Refinements" for n in names)
    for t in range(3):
        assert np.array_equal(arrays[f"{names[t]}Refinements"], it.getRefinementsRequired(t)), t
    # fixed level 1: original and refined meshes to OBJ and VTK
    run_fixed = tmp_path / "fixed"
    run_fixed.mkdir()
    out = run_driver(["-f", str(dat), "-r", "1", "--exporttoobj", "--exporttovtk", "--exportresults"], str(run_fixed))
    assert "Using fixed refinement level equal to 1" in out
    V, T = len(xyz), len(tri)
    obj = open(run_fixed / "RefinedMesh.obj").read().split("\n")[:-1]
    assert sum(l.startswith("v ") for l in obj) == V + 3 * T and sum(l.startswith("f ") for l in obj) == 4 * T
    obj0 = open(run_fixed / "OriginalMesh.obj").read().split("\n")[:-1]
    assert len(obj0) == V + T
    root = ET.fromstring(open(run_fixed / "RefinedMesh.vtp").read().split("\n", 1)[1])
    piece = root.find("PolyData/Piece")
    assert piece.get("NumberOfPoints") == str(V + 3 * T) and piece.find("CellData") is None
    assert os.path.exists(run_fixed / "OriginalMesh.vtp")
    assert all(os.path.exists(run_fixed / f"{n}.dat") for n in names)


def brute_force_lists(tri):
    """kDetermineNeighborType's definition (src/Mesh3d.cu:93-142): for i < j,
    the number of cell i's three vertex slots that occur in cell j."""
    T = len(tri)
    simple, attached, disjoint = [], [], []
    for i in range(T):
        for j in range(i + 1, T):
            m = sum(int(tri[i][k] in tri[j]) for k in range(3))
            (disjoint if m == 0 else simple if m == 1 else attached if m == 2 else []).append((i, j))
    return simple, attached, disjoint


def test_classification_on_device(gpu):
    """The incidence-merge classification equals the brute-force definition,
    including repeated cells, cells with a repeated vertex and a vertex of
    valence 40; the counts at full size equal the icosphere's closed forms."""
    rng = np.random.default_rng(7)
    xyz, tri = gpu.icosphere(1, (1.0, 1.0, 1.0), 0.0, 0)
    V = len(xyz)
    fan = np.array([[V, 1 + k % (V - 1), 1 + (k + 1) % (V - 1)] for k in range(40)], np.int32)
    extra = np.array([tri[5], [tri[7][0], tri[7][0], tri[7][1]], [3, 3, 3]], np.int32)
    tri2 = np.concatenate([tri, fan, extra])[rng.permutation(len(tri) + 43)]
    xyz2 = np.concatenate([xyz, [[0.0, 0.0, 2.0]]])
    m = gpu.Mesh3D()
    m.setData(xyz2, tri2)
    m.prepareMesh()
    s, a, d = brute_force_lists([list(t) for t in tri2])
    k = lambda l: np.array([(i, j, q) for q, (i, j) in enumerate(l)], np.int32).reshape(-1, 3)
    assert np.array_equal(m.getSimpleNeighbors(), k(s))
    assert np.array_equal(m.getAttachedNeighbors(), k(a))
    assert np.array_equal(m.getNotNeighbors(), k(d))
    assert m.sizes()[2] == [len(s), len(a), len(d)]
    for lev in (4, 5):
        xyz, tri = gpu.icosphere(lev)
        m = gpu.Mesh3D()
        m.setData(xyz, tri)
        m.prepareMesh()
        V, T = len(xyz), len(tri)
        vertex, edge = 60 + 9 * (V - 12), 3 * T // 2
        assert m.sizes()[2] == [vertex, edge, T * (T - 1) // 2 - vertex - edge], lev
