"""Benchmark of the panel-pair integrator on one MI355X.

Workload (one "step"): all 26,209,280 ordered panel pairs of the jittered
ellipsoid -- icosphere subdivided 4x (2,562 vertices, 5,120 triangles), semi-axes
(3, 1, 0.5), vertices moved along the normal by N(0, (0.05 * mean edge)^2) from
seed 0 -- integrated in FP64 with the 13-point rule and automatic Runge error
control (up to 3 refinement levels, tolerance 1e-8), all three neighbour types,
analytic singular parts and the final J, inputs resident in HBM.

Prints one JSON line: value = ordered pair-integrals per second of the whole
step.  `roofline` is for the dominant kernel (disjoint pairs, refinement level 1),
timed with HIP events on the library's stream; `cpu_baseline` times the CPU
oracle (FP64 OpenMP restatement) on a stratified sample of the same tasks.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import integrator2_amd as i2  # noqa: E402

N_PAIRS_CONFIG2 = 26_209_280
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic FP64 work of one quadrature point of the inner integral
# (thetaPsi, evaluatorJ3DK.cu:266-313, with the edge tangents recomputed per
# point as written, plus the Gauss point and the weighted sum): 188 basic
# operations, 6 square roots, 3 logarithms and 1 atan2.  Transcendentals are
# weighted by their gfx950 FP64 VALU instruction counts (TRANS_WEIGHTS).
BASIC_OPS_PER_POINT = 188
TRANS_PER_POINT = {"sqrt": 6, "log": 3, "atan2": 1}
TRANS_WEIGHTS = {"sqrt": 8, "log": 22, "atan2": 40}
POINTS = 13


def flops_per_point():
    return BASIC_OPS_PER_POINT + sum(TRANS_PER_POINT[k] * TRANS_WEIGHTS[k] for k in TRANS_PER_POINT)


def make_config2():
    return i2.icosphere(4, (3.0, 1.0, 0.5), 0.05, 0)


def run_plain(ev):
    ev.runAllPairs(False)


def cpu_baseline(xyz, tri, conv_by_type, budget_s=20.0):
    """Times the CPU oracle on a stratified sample of the step's ordered tasks
    (all refinement included) and extrapolates to the full task mix."""
    import oracle
    orc = oracle.Oracle(xyz, tri)
    threads = oracle.lib().o_max_threads()
    rng = np.random.default_rng(0)
    total_s = 0.0
    sample_desc = []
    for t in range(3):
        tasks = orc.ordered_tasks(t)
        n = len(tasks)
        m = min(n, 4000 if t < 2 else 20000)
        idx = rng.choice(n, size=m, replace=False) if m < n else np.arange(n)
        t0 = time.perf_counter()
        orc.run_tasks(t, tasks[idx], i2.QF3D13_L12, i2.QF3D13_W, 7, mode=1, level=3, tol=1e-8)
        dt = time.perf_counter() - t0
        total_s += dt * n / m
        sample_desc.append(f"{m}/{n} {i2.TYPE_NAMES[t]}")
    return {"value": N_PAIRS_CONFIG2 / total_s, "unit": "pair-integrals/s", "cores": threads, "kind": "port",
            "sample": "stratified random ordered tasks, full adaptive refinement: " + ", ".join(sample_desc)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if i2.device_count() < 1:
        raise SystemExit("no HIP device")

    xyz, tri = make_config2()
    mesh = i2.Mesh3D(0)
    mesh.setData(xyz, tri)
    mesh.prepareMesh()
    integ = i2.NumericalIntegrator3D(mesh)
    integ.setErrorControl(3, 1e-8, True)
    ev = i2.EvaluatorJ3DK(mesh, integ)

    for _ in range(args.warmup):
        run_plain(ev)
    i2.synchronize(0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_plain(ev)
    i2.synchronize(0)
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = N_PAIRS_CONFIG2 / (elapsed / args.steps)

    # dominant kernel: disjoint pairs at refinement level 1, timed with events
    ev.setProfiling(True)
    run_plain(ev)
    recs = ev.launchRecords()
    ev.setProfiling(False)
    dom = max(recs, key=lambda r: r[3])
    dom_type, dom_level, dom_tasks, dom_ms = dom
    sub_panels = 4 ** dom_level
    flops = dom_tasks * sub_panels * POINTS * flops_per_point()
    achieved_tflops = flops / (dom_ms * 1e-3) / 1e12
    # HBM traffic per launch (algorithmic): one double4 written per task, 192-byte
    # panel records of both panels read once
    traffic = dom_tasks * 32 + 2 * tri.shape[0] * 192
    refined = {t: ev.typeStats(t)[1] for t in range(3)}
    Ks = {t: ev.refinementIterations(t) for t in range(3)}

    out = {
        "metric": "FP64 pair-integrals/s (<=1e-11 rel err) batched vs plain; adjoint/forward ratio",
        "value": value,
        "unit": "pair-integrals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "jittered ellipsoid (3,1,0.5), icosphere-4: 5120 triangles, 26,209,280 ordered pairs, "
                               "13-point rule, auto refinement <=3 levels at tol 1e-8",
                   "path": "plain (per-type, per-level launches)",
                   "refinement_iterations": Ks, "refined_tasks": refined},
        "roofline": {"bound": "valu-fp64", "achieved": achieved_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": f"k_level type {dom_type} level {dom_level}", "kernel_ms": dom_ms,
                     "algorithmic_bytes": traffic, "flops_per_point": flops_per_point()},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(xyz, tri, None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
