"""Shared helpers of the test-suite: workload meshes and parity metrics."""
import numpy as np

TOL = 1e-11   # norm-wise relative error per ordered pair


def rel_err_rows(gpu, cpu):
    """Norm-wise relative error per row: ||gpu - cpu||_inf / ||cpu||_inf."""
    gpu = np.asarray(gpu, float)
    cpu = np.asarray(cpu, float)
    num = np.abs(gpu - cpu).max(axis=1)
    den = np.abs(cpu).max(axis=1)
    return num / np.maximum(den, 1e-300)


def reported_levels(conv, K, quirk, fixed=False):
    """Level whose value a task reports: its convergence iteration c, or c-1
    when the reference's buffer swap applies (K - c odd)."""
    conv = np.asarray(conv, np.int64)
    if fixed or not quirk:
        return conv.astype(np.int8)
    r = np.where((conv >= 1) & (((K - conv) & 1) == 1), conv - 1, conv)
    return r.astype(np.int8)


def parity_gate(orc, i2, t, tasks, gpu, cpu, rlevel, label, tol=TOL, results=False):
    """1e-11 norm-wise relative error per ordered pair; a task above it must lie
    within 1e-11 + 4*nu, nu being the FP64 rounding noise of the reference's
    formulas for that task (oracle in double vs long double)."""
    e = rel_err_rows(gpu, cpu)
    bad = np.nonzero(e > tol)[0]
    nu = np.zeros(len(e))
    if len(bad):
        nu[bad] = orc.task_noise(t, np.asarray(tasks)[bad], np.asarray(rlevel)[bad], i2.QF3D13_L12, i2.QF3D13_W,
                                 i2.QF3D13_ORDER, results=results)
    bound = tol + 4.0 * nu
    ratio = (e / bound).max() if len(e) else 0.0
    print(f"{label}: n {len(e)}, max rel err {e.max() if len(e) else 0:.3e}, above 1e-11: {len(bad)} "
          f"(max noise {nu.max() if len(e) else 0:.2e}), worst err/bound {ratio:.3f}")
    assert (e <= bound).all(), f"{label}: {(e > bound).sum()} tasks beyond the gate"
    return e, nu


def config2_mesh(i2, level=4):
    """Icosphere-4 mapped to the (3, 1, 0.5) ellipsoid with 0.05-edge normal jitter (seed 0)."""
    return i2.icosphere(level, (3.0, 1.0, 0.5), 0.05, 0)
