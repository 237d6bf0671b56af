"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``oracle/build/liboracle.so`` (oracle.cpp), the CPU
restatement of the reference solver path. Only tests/, the smoke entry point and
the CPU-baseline leg of bench.py may import this module; it is the checker,
never the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_L = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_ex_create.restype = C.c_void_p
        L.orc_ex_create.argtypes = [C.c_void_p]
        L.orc_ex_update.argtypes = [C.c_void_p, C.c_int]
        L.orc_ex_calculate.argtypes = [C.c_void_p]
        L.orc_ex_snapshot.argtypes = [C.c_void_p, _dp]
        L.orc_ex_gates.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_ex_destroy.argtypes = [C.c_void_p]
        L.orc_im_create.restype = C.c_void_p
        L.orc_im_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                    C.c_int, C.c_int, C.c_int]
        L.orc_im_update.argtypes = [C.c_void_p, C.c_int]
        L.orc_im_calculate.argtypes = [C.c_void_p]
        L.orc_im_snapshot.argtypes = [C.c_void_p, _dp]
        L.orc_im_matrix.argtypes = [C.c_void_p, _ip, _ip, _dp, _dp]
        L.orc_im_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int), _dp, C.POINTER(C.c_int64)]
        L.orc_im_destroy.argtypes = [C.c_void_p]
        L.orc_tree_solve.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.orc_run.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                              C.c_int, C.c_int64, _dp, C.POINTER(C.c_int64)]
        L.orc_setup.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_active_range.argtypes = [C.c_double, C.c_int64, C.c_double, C.c_double,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _L = L
    return _L


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Explicit:
    """solverEx-cpu.cpp semantics."""

    def __init__(self, problem):
        self._s = problem.struct()
        self._p = problem
        self.h = lib().orc_ex_create(C.byref(self._s))
        self.N = problem.N

    def update(self, active):
        lib().orc_ex_update(self.h, int(bool(active)))

    def calculate(self):
        lib().orc_ex_calculate(self.h)

    def snapshot(self):
        out = np.empty(self.N)
        lib().orc_ex_snapshot(self.h, out.ctypes.data_as(_dp))
        return out

    def gates(self):
        H = self._p.H
        m, h, n = np.empty(H), np.empty(H), np.empty(H)
        lib().orc_ex_gates(self.h, m.ctypes.data_as(_dp), h.ctypes.data_as(_dp),
                           n.ctypes.data_as(_dp))
        return m, h, n

    def run(self, first, num, a0=-1, a1=-2, trace_elem=0, interval=-1):
        cap = num // (interval + 1) if interval >= 0 else 0
        tr = np.empty(max(cap, 1))
        ns = C.c_int64(0)
        st = lib().orc_run(self.h, 0, first, num, a0, a1, trace_elem, interval,
                           tr.ctypes.data_as(_dp), C.byref(ns))
        return tr[:ns.value], st

    def close(self):
        if self.h:
            lib().orc_ex_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


VAR_CPU, VAR_GPU = 0, 1
M_GMRES, M_BICGSTAB, M_DIRECT = 0, 1, 2


class Implicit:
    """solverIm-cpu.cpp (variant CPU) / solverIm-gpu.cpp (variant GPU) semantics."""

    def __init__(self, problem, variant=VAR_GPU, method=M_BICGSTAB, restart=150, tol=1e-14,
                 tol_mode=0, x0_mode=1, check_every=30, max_iter=0):
        self._s = problem.struct()
        self._p = problem
        self.h = lib().orc_im_create(C.byref(self._s), variant, method, restart, tol, tol_mode,
                                     x0_mode, check_every, max_iter)
        self.N = problem.N

    def update(self, active):
        lib().orc_im_update(self.h, int(bool(active)))

    def calculate(self):
        return lib().orc_im_calculate(self.h)

    def snapshot(self):
        out = np.empty(self.N)
        lib().orc_im_snapshot(self.h, out.ctypes.data_as(_dp))
        return out

    def matrix(self):
        N = self.N
        rowptr = np.empty(N + 1, dtype=np.int32)
        nnz = lib().orc_im_matrix(self.h, rowptr.ctypes.data_as(_ip), None, None, None)
        col = np.empty(nnz, dtype=np.int32)
        val = np.empty(nnz)
        rhs = np.empty(N)
        lib().orc_im_matrix(self.h, rowptr.ctypes.data_as(_ip), col.ctypes.data_as(_ip),
                            val.ctypes.data_as(_dp), rhs.ctypes.data_as(_dp))
        return rowptr, col, val, rhs

    def stats(self):
        it, res, tot = C.c_int(0), C.c_double(0), C.c_int64(0)
        lib().orc_im_stats(self.h, C.byref(it), C.byref(res), C.byref(tot))
        return it.value, res.value, tot.value

    def run(self, first, num, a0=-1, a1=-2, trace_elem=0, interval=-1):
        cap = num // (interval + 1) if interval >= 0 else 0
        tr = np.empty(max(cap, 1))
        ns = C.c_int64(0)
        st = lib().orc_run(self.h, 1, first, num, a0, a1, trace_elem, interval,
                           tr.ctypes.data_as(_dp), C.byref(ns))
        return tr[:ns.value], st

    def close(self):
        if self.h:
            lib().orc_im_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def tree_solve(rowptr, col, val, rhs):
    N = len(rhs)
    x = np.empty(N)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val, rhs = _d(val), _d(rhs)
    st = lib().orc_tree_solve(N, rowptr.ctypes.data_as(_ip), col.ctypes.data_as(_ip),
                              val.ctypes.data_as(_dp), rhs.ctypes.data_as(_dp),
                              x.ctypes.data_as(_dp))
    if st != 0:
        raise ValueError("matrix is not a tree system")
    return x


def setup(geom_struct, out_struct):
    """orc_setup on the C structs of neurite_hip (GeometryStruct, SetupOutStruct)."""
    return lib().orc_setup(C.byref(geom_struct), C.byref(out_struct))


def active_range(dT, num_steps_T, init_time, end_input):
    a0, a1 = C.c_int64(0), C.c_int64(0)
    lib().orc_active_range(dT, num_steps_T, init_time, end_input, C.byref(a0), C.byref(a1))
    return a0.value, a1.value
