"""Edge cases of the device paths: a mesh without any pair (one triangle),
and a type beyond the plain path's 32-bit task lists (the 65,536-triangle
torus has 4.29e9 ordered disjoint pairs), which must fail with a clear error
instead of overflowing as the reference's int counters would."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path", [0, 1])
def test_mesh_without_pairs(gpu, tmp_path, path):
    xyz = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    tri = np.array([[0, 1, 2]], np.int32)
    m = gpu.Mesh3D()
    m.setData(xyz, tri)
    m.prepareMesh()
    assert m.sizes()[2] == [0, 0, 0]
    it = gpu.NumericalIntegrator3D(m)
    ev = gpu.EvaluatorJ3DK(m, it)
    ev.setPath(path)
    for fixed in (None, 2):
        if fixed is None:
            it.setErrorControl(3, 1e-8)
        else:
            it.setFixedRefinementLevel(fixed)
        print(f"path {path}, {'automatic' if fixed is None else 'fixed level 2'}", flush=True)
        ev.runAllPairs(True)
        for t in range(3):
            r = ev.results(t, errors=True)
            assert len(r["tasks"]) == 0 and len(r["integrals"]) == 0 and len(r["errors"]) == 0
            if fixed is None:
                assert ev.refinementIterations(t) == 0
            # the reference's outputResultsToFile returns false for a type without tasks
            assert not ev.outputResultsToFile(t, gpu.CSV, str(tmp_path))
            counts = it.getRefinementsRequired(t)
            assert (len(counts) == 0) if fixed is not None else np.array_equal(counts, [0])
    assert os.listdir(tmp_path) == []


def test_plain_path_refuses_more_than_32_bit_tasks(gpu):
    xyz, tri = gpu.torus(256, 128, 1.0, 0.25, 0.02, 0)
    m = gpu.Mesh3D()
    m.setData(xyz, tri)
    m.prepareMesh()
    assert 2 * m.sizes()[2][2] > 2 ** 31 - 1
    it = gpu.NumericalIntegrator3D(m)
    it.setFixedRefinementLevel(0)
    ev = gpu.EvaluatorJ3DK(m, it)
    ev.setPath(0)
    with pytest.raises(gpu.I2Error, match="32-bit"):
        ev.runAllPairs(False)


def test_results_follow_the_path_of_the_last_run(gpu):
    """Changing the path after a run does not change what the run's getters read."""
    xyz, tri = gpu.icosphere(1, (2.0, 1.0, 0.7), 0.05, 2)
    m = gpu.Mesh3D()
    m.setData(xyz, tri)
    m.prepareMesh()
    it = gpu.NumericalIntegrator3D(m)
    it.setErrorControl(2, 1e-8)
    ev = gpu.EvaluatorJ3DK(m, it)
    ev.setPath(1)
    ev.runAllPairs(True)
    batched = [ev.results(t, errors=True) for t in range(3)]
    counts = [it.getRefinementsRequired(t) for t in range(3)]
    ev.setPath(0)
    for t in range(3):
        r = ev.results(t, errors=True)
        for key in ("tasks", "integrals", "results", "conv", "errors"):
            assert np.array_equal(r[key], batched[t][key]), (t, key)
        assert np.array_equal(it.getRefinementsRequired(t), counts[t])
    ev.runAllPairs(True)
    for t in range(3):
        r = ev.results(t)
        assert np.array_equal(r["conv"], batched[t]["conv"])
        e = np.abs(r["integrals"] - batched[t]["integrals"]).max(axis=1) / np.abs(batched[t]["integrals"]).max(axis=1)
        assert e.max() <= 1e-11, (t, e.max())
