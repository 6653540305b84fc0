"""The reference's other triangle rules (qf3D1 ... qf3D12) on the device:
fixed level 1 and automatic refinement, plain and batched paths, against the
oracle run with the same rule (identical refinement decisions, 1e-11
norm-wise, noise-aware gate) and the symmetry delta of the returned J."""
import numpy as np
import pytest

from tests.helpers import parity_gate, reported_levels

pytestmark = pytest.mark.gpu


def run(i2, xyz, tri, rule, path, fixed=None, maxLevel=2, tol=1e-6):
    L12, W, order = rule
    m = i2.Mesh3D()
    m.setData(xyz, tri)
    m.prepareMesh()
    it = i2.NumericalIntegrator3D(m, L12.ravel(), W, order)
    if fixed is not None:
        it.setFixedRefinementLevel(fixed)
    else:
        it.setErrorControl(maxLevel, tol, True)
    ev = i2.EvaluatorJ3DK(m, it)
    ev.setPath(path)
    ev.runAllPairs(True)
    return {t: ev.results(t, errors=True) for t in range(3)}, {t: ev.refinementIterations(t) for t in range(3)}


@pytest.mark.parametrize("npts", [1, 3, 4, 6, 7, 9, 12])
def test_rule_on_device(gpu, oracle_mod, npts):
    rule = gpu.quadrature_rule(npts)
    L12, W, order = rule
    xyz, tri = gpu.icosphere(2, (2.0, 1.0, 0.7), 0.05, 0)
    orc = oracle_mod.Oracle(xyz, tri)
    for path in (0, 1):
        for fixed in (1, None):
            out, Ks = run(gpu, xyz, tri, rule, path, fixed=fixed)
            for t in range(3):
                r = out[t]
                if fixed is not None:
                    I4, J3, conv, K = orc.run_tasks(t, r["tasks"], L12.ravel(), W, order, mode=0, level=fixed)
                    rl = np.full(len(conv), fixed, np.int8)
                else:
                    I4, J3, conv, K = orc.run_tasks(t, r["tasks"], L12.ravel(), W, order, mode=1, level=2, tol=1e-6)
                    assert K == Ks[t]
                    assert np.array_equal(conv, r["conv"]), f"{npts} points path {path} type {t}: decisions differ"
                    rl = reported_levels(conv, K, True)
                label = f"{npts} points path {path} {'L1' if fixed else 'auto'} type {t}"
                parity_gate(orc, gpu, t, r["tasks"], r["integrals"], I4, rl, label + " integrals", rule=(L12.ravel(), W, order))
                parity_gate(orc, gpu, t, r["tasks"], r["results"], J3, rl, label + " results", results=True,
                            rule=(L12.ravel(), W, order))
                # the symmetry delta is the reference's formula on the returned J (the
                # values themselves are gated above, noise-aware)
                ref = oracle_mod.symmetry(r["results"])
                assert np.abs(r["errors"] - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
