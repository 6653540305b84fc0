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


def runge_criterion(cur, prev, pow2p=128.0):
    """Runge criterion of the reference (evaluator3d.cu:83-92 with the divide
    guards of cuda_math.cuh:270-280), vectorised, for classifying decisions."""
    num = cur - prev
    den = pow2p * prev - cur
    out = np.zeros(len(cur))
    for c in range(4):
        eps = 1e-10 if c == 3 else 2e-6
        small = (np.abs(num[:, c]) < eps) & (np.abs(den[:, c]) < eps)
        with np.errstate(divide="ignore", invalid="ignore"):
            out += np.where(small, 0.0, np.abs(num[:, c] / den[:, c]))
    return out


def check_decisions(conv_gpu, conv_orc, levels_orc, tol, label, margin=1e-2, max_frac=1e-5):
    """Refinement decisions must agree except for tasks whose Runge criterion
    at the first diverging level lies within `margin` of the tolerance: there
    the 1e-11 parity noise of the level values (divided by their ~1e-8
    difference) can move the criterion across it.  Returns the flipped indices."""
    conv_gpu = np.asarray(conv_gpu, np.int64)
    conv_orc = np.asarray(conv_orc, np.int64)
    flips = np.nonzero(conv_gpu != conv_orc)[0]
    if len(flips):
        L = np.minimum(conv_gpu[flips], conv_orc[flips])
        crit = runge_criterion(levels_orc[flips, L], levels_orc[flips, L - 1])
        rel = np.abs(crit - tol) / tol
        print(f"{label}: {len(flips)} refinement decisions differ of {len(conv_gpu)}, "
              f"criterion within {rel.max():.1e} of the tolerance", flush=True)
        assert (L >= 1).all() and (rel <= margin).all(), (label, rel.max())
    assert len(flips) <= max(2, max_frac * len(conv_gpu)), (label, len(flips))
    return flips
