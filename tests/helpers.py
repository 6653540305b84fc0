"""Shared helpers of the test-suite: workload meshes and parity metrics."""
import numpy as np


def rel_err_rows(gpu, cpu):
    """Norm-wise relative error per row: ||gpu - cpu||_inf / ||cpu||_inf."""
    gpu = np.asarray(gpu, float)
    cpu = np.asarray(cpu, float)
    num = np.abs(gpu - cpu).max(axis=1)
    den = np.abs(cpu).max(axis=1)
    return num / np.maximum(den, 1e-300)


def icosphere_mesh(i2, level, axes=(1.0, 1.0, 1.0), jitter=0.0, seed=0):
    return i2.icosphere(level, axes, jitter, seed)


def config2_mesh(i2, level=4):
    """Icosphere-4 mapped to the (3, 1, 0.5) ellipsoid with 0.05-edge normal jitter (seed 0)."""
    return i2.icosphere(level, (3.0, 1.0, 0.5), 0.05, 0)
