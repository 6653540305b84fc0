"""ctypes wrapper of the CPU oracle (test infrastructure only).

The oracle is the checker of the HIP product path: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.o_create.restype = _P
        L.o_pair_count.restype = ctypes.c_int64
        L.o_normal_draw.restype = ctypes.c_double
        L.o_normal_draw.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        L.o_loss_tasks.restype = ctypes.c_double
        L.o_loss_tasks_prec.restype = ctypes.c_double
        L.o_now.restype = ctypes.c_double
        L.o_singular_integral.restype = ctypes.c_uint
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


class Oracle:
    """CPU restatement of the reference for one mesh."""

    def __init__(self, xyz, tri):
        self.xyz = np.ascontiguousarray(xyz, np.float64)
        self.tri = np.ascontiguousarray(tri, np.int32)
        self._h = _P(lib().o_create(_ptr(self.xyz), _I64(len(self.xyz)), _ptr(self.tri), _I64(len(self.tri))))
        self.nT = len(self.tri)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.o_destroy(self._h)
            self._h = None

    def pair_count(self, t):
        return lib().o_pair_count(self._h, ctypes.c_int(t))

    def neighbor_list(self, t):
        n = self.pair_count(t)
        out = np.zeros((n, 3), np.int32)
        lib().o_get_list(self._h, ctypes.c_int(t), _ptr(out))
        return out

    def ordered_tasks(self, t):
        l = self.neighbor_list(t)
        return np.ascontiguousarray(np.concatenate([l[:, :2], l[:, [1, 0]]]).astype(np.int32))

    def normals(self):
        out = np.zeros((self.nT, 3))
        lib().o_get_normals(self._h, _ptr(out))
        return out

    def measures(self):
        out = np.zeros(self.nT)
        lib().o_get_measures(self._h, _ptr(out))
        return out

    def run_tasks(self, t, ij, L12, w, order, mode=1, level=5, tol=1e-5, quirk=True, K_override=-1,
                  nthreads=0, want_levels=False, long_double=False, want_unconv=False):
        """Integrals (n,4), J (n,3), conv (n,), K [, levels (n, level+1, 4)] [, unconv (n,)] for
        ordered tasks ij (n,2).

        long_double=True evaluates the level values and the singular part in
        long double (each rounded once to double) before the Runge decisions.
        unconv[q] = 1 where task q was unconverged after its last comparison."""
        ij = np.ascontiguousarray(np.asarray(ij)[:, :2], np.int32)
        n = len(ij)
        L12 = np.ascontiguousarray(L12, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        I4 = np.zeros((n, 4))
        J3 = np.zeros((n, 3))
        conv = np.zeros(n, np.int8)
        lev = np.zeros((n, level + 1, 4)) if want_levels else None
        K = ctypes.c_int()
        unconv = np.zeros(n, np.uint8) if want_unconv else None
        lib().o_run_tasks_prec(self._h, ctypes.c_int(t), _ptr(ij), _I64(n), _ptr(L12), _ptr(w), ctypes.c_int(len(w)),
                               ctypes.c_int(order), ctypes.c_int(mode), ctypes.c_int(level), _D(tol),
                               ctypes.c_int(int(quirk)), ctypes.c_int(K_override), ctypes.c_int(nthreads),
                               ctypes.c_int(int(long_double)), _ptr(I4), _ptr(J3), _ptr(conv), _ptr(lev), ctypes.byref(K),
                               _ptr(unconv))
        out = (I4, J3, conv, K.value) + ((lev,) if want_levels else ()) + ((unconv,) if want_unconv else ())
        return out

    def refinement_counts(self, ij, conv, unconv, max_level):
        """Per-cell refinement counts of one type from its tasks' conv / unconv
        (NumericalIntegrator3D::getRefinementsRequired)."""
        ij = np.ascontiguousarray(np.asarray(ij)[:, :2], np.int32)
        conv = np.ascontiguousarray(conv, np.int8)
        unconv = np.ascontiguousarray(unconv, np.uint8)
        out = np.zeros(len(self.tri), np.uint8)
        lib().o_refinement_counts(ctypes.c_int(len(self.tri)), _ptr(ij), _I64(len(ij)), _ptr(conv), _ptr(unconv),
                                  ctypes.c_int(max_level), _ptr(out))
        return out

    def theta_psi(self, pts, tri9):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        tri9 = np.ascontiguousarray(tri9, np.float64).reshape(9)
        out = np.zeros((len(pts), 4))
        lib().o_theta_psi(_ptr(pts), _I64(len(pts)), _ptr(tri9), _ptr(out))
        return out

    def singular_part(self, t, i, j, pts):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros((len(pts), 4))
        lib().o_singular_part(self._h, ctypes.c_int(t), ctypes.c_int(i), ctypes.c_int(j), _ptr(pts), _I64(len(pts)), _ptr(out))
        return out

    def singular_integral(self, t, i, j, with_branches=False):
        """Closed-form singular part of task (i, j) [, special-case branch bit mask]."""
        out = np.zeros(4)
        br = lib().o_singular_integral(self._h, ctypes.c_int(t), ctypes.c_int(i), ctypes.c_int(j), _ptr(out))
        return (out, br) if with_branches else out

    def singular_integrals(self, t, ij):
        """Closed-form singular parts of ordered tasks in double and in long double."""
        ij = np.ascontiguousarray(np.asarray(ij)[:, :2], np.int32)
        d = np.zeros((len(ij), 4))
        ld = np.zeros((len(ij), 4))
        lib().o_singular_integrals(self._h, ctypes.c_int(t), _ptr(ij), _I64(len(ij)), _ptr(d), _ptr(ld))
        return d, ld

    def fp64_noise(self, t, ij, totals):
        """Per-task FP64 rounding noise of the reference's closed forms, relative to the task's total."""
        if t == 2:
            return np.zeros(len(ij))
        d, ld = self.singular_integrals(t, ij)
        return np.abs(d - ld).max(axis=1) / np.maximum(np.abs(np.asarray(totals)).max(axis=1), 1e-300)

    def values_at_levels(self, t, ij, rlevel, L12, w, order, long_double=False):
        """Integrals (n,4) and J (n,3) of ordered tasks at the given reported levels
        (evaluated in long double when long_double)."""
        ij = np.ascontiguousarray(np.asarray(ij)[:, :2], np.int32)
        rlevel = np.ascontiguousarray(rlevel, np.int8)
        L12 = np.ascontiguousarray(L12, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        I4 = np.zeros((len(ij), 4))
        J3 = np.zeros((len(ij), 3))
        lib().o_tasks_at_levels_prec(self._h, ctypes.c_int(t), _ptr(ij), _ptr(rlevel), _I64(len(ij)), _ptr(L12),
                                     _ptr(w), ctypes.c_int(len(w)), ctypes.c_int(order), ctypes.c_int(int(long_double)),
                                     _ptr(I4), _ptr(J3))
        return I4, J3

    def level_value(self, t, i, j, L12, w, order, level):
        L12 = np.ascontiguousarray(L12, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        out = np.zeros(4)
        lib().o_level_value(self._h, ctypes.c_int(t), ctypes.c_int(i), ctypes.c_int(j), _ptr(L12), _ptr(w),
                            ctypes.c_int(len(w)), ctypes.c_int(order), ctypes.c_int(level), _ptr(out))
        return out

    def loss_tasks(self, t, ij, rlevel, L12, w, order, seed, vtx=-1, axis=0, h=0.0, dual=False, nthreads=0,
                   long_double=False):
        ij = np.ascontiguousarray(np.asarray(ij)[:, :2], np.int32)
        rlevel = np.ascontiguousarray(rlevel, np.int8)
        L12 = np.ascontiguousarray(L12, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        return lib().o_loss_tasks_prec(self._h, ctypes.c_int(t), _ptr(ij), _ptr(rlevel), _I64(len(ij)), _ptr(L12),
                                       _ptr(w), ctypes.c_int(len(w)), ctypes.c_int(order), ctypes.c_uint64(seed),
                                       ctypes.c_int(vtx), ctypes.c_int(axis), _D(h), ctypes.c_int(int(dual)),
                                       ctypes.c_int(int(long_double)), ctypes.c_int(nthreads))


def symmetry(results3):
    r = np.ascontiguousarray(results3, np.float64)
    n = len(r) // 2
    out = np.zeros(2 * n)
    lib().o_symmetry(_ptr(r), _I64(n), _ptr(out))
    return out


def normal_draw(seed, pair_id, stream):
    return lib().o_normal_draw(seed, pair_id, stream)


def _task_noise(self, t, ij, rlevel, L12, w, order, results=False):
    """Per-task FP64 rounding noise of the reference's formulas (double vs long
    double), relative to the task's integrals, or to its J when results=True."""
    ij = np.ascontiguousarray(np.asarray(ij)[:, :2], np.int32)
    rlevel = np.ascontiguousarray(rlevel, np.int8)
    L12 = np.ascontiguousarray(L12, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    out = np.zeros(len(ij))
    out_j = np.zeros(len(ij))
    lib().o_task_noise(self._h, ctypes.c_int(t), _ptr(ij), _ptr(rlevel), _I64(len(ij)), _ptr(L12), _ptr(w),
                       ctypes.c_int(len(w)), ctypes.c_int(order), _ptr(out), _ptr(out_j))
    return out_j if results else out


Oracle.task_noise = _task_noise


def _task_noise_mca(self, t, ij, rlevel, L12, w, order, results=False, nrep=3):
    """Monte Carlo arithmetic noise (every operation randomly perturbed by up to
    one ulp) of tasks at their reported levels, relative to the integrals or J."""
    ij = np.ascontiguousarray(np.asarray(ij)[:, :2], np.int32)
    rlevel = np.ascontiguousarray(rlevel, np.int8)
    L12 = np.ascontiguousarray(L12, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    out = np.zeros(len(ij))
    out_j = np.zeros(len(ij))
    lib().o_task_noise_mca(self._h, ctypes.c_int(t), _ptr(ij), _ptr(rlevel), _I64(len(ij)), _ptr(L12), _ptr(w),
                           ctypes.c_int(len(w)), ctypes.c_int(order), ctypes.c_int(nrep), _ptr(out), _ptr(out_j))
    return out_j if results else out


Oracle.task_noise_mca = _task_noise_mca
