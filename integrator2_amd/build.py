"""Build recipe for the gfx950 shared library ``libintegrator2_mi355x.so`` and
the command-line driver ``integrator2test3D``.

hipcc compiles every HIP/C++ source of ``csrc/`` for gfx950 only; the
library is written in-tree next to this file so that it travels with the
repository snapshot.  The driver (``cpp/``) is host C++ on the reference's
class API (``include/integrator2/integrator2.hpp``), built with g++ against
the library.  ``build()`` always recompiles both.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libintegrator2_mi355x.so")
DRIVER = os.path.join(HERE, "integrator2test3D")

SOURCES = ["i2_mesh.hip", "i2_evaluator.hip", "i2_batch.hip", "i2_adjoint.hip", "i2_refine.hip", "i2_meshgen.cpp",
           "i2_rules_io.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # keep hand-written FP32 packed math out of the code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if s.endswith(".hip") or s.endswith(".cpp")]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


def build_driver(verbose=False):
    """The reference-style driver: g++ on the C++ API header, linked to the library."""
    tmp = DRIVER + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", "/opt/rocm/include", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "cpp", "integrator2test3D.cpp"),
           "-L", HERE, "-lintegrator2_mi355x", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
    sys.exit(0)
