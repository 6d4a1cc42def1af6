"""ctypes front end of the OpenMP CPU baseline (libjfem_cpu.so built from cpu_baseline.c).

BENCHMARK AND TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg and by tests/,
never by the product.  It restates the reference's CPU solve path (row-parallel SparseMvMul
products, mgStepV, solvePCG) on all host cores, with the level systems passed in as host CSR
arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import platform

import numpy as np

from . import oracle as _O

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libjfem_cpu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            _O.build()
        L = C.CDLL(_LIB_PATH)
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        L.cb_threads.restype = C.c_int
        L.cb_mg_create.restype = C.c_void_p
        L.cb_mg_create.argtypes = [C.c_int, C.c_int, ip, C.c_int, lp, C.POINTER(lp), C.POINTER(ip), C.POINTER(dp),
                                   C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]
        L.cb_mg_free.argtypes = [C.c_void_p]
        L.cb_mg_prol_nnz.restype = C.c_long
        L.cb_mg_prol_nnz.argtypes = [C.c_void_p, C.c_int]
        L.cb_mg_prol_csr.argtypes = [C.c_void_p, C.c_int, lp, ip, dp]
        L.cb_mg_lmax.restype = C.c_double
        L.cb_mg_lmax.argtypes = [C.c_void_p, C.c_int]
        L.cb_mg_vcycle.argtypes = [C.c_void_p, dp, dp]
        L.cb_mg_pcg.restype = C.c_int
        L.cb_mg_pcg.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_long, dp, dp]
        L.cb_mg_spmv_seconds.restype = C.c_double
        L.cb_mg_spmv_seconds.argtypes = [C.c_void_p, C.c_int]
        _lib = L
    return _lib


def threads() -> int:
    return lib().cb_threads()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


class CpuMG:
    """Chebyshev-Jacobi V-cycle hierarchy over host CSR level systems (coarsest first)."""

    def __init__(self, d, k, coarse_cells, levels, degree=4, runs=2, ratio=20.0, safety=1.1, power_its=30):
        """levels: (rowptr int64, cols int32, vals float64) per level, coarsest first"""
        self._keep = [(np.ascontiguousarray(rp, np.int64), np.ascontiguousarray(c, np.int32),
                       np.ascontiguousarray(v, np.float64)) for rp, c, v in levels]
        nl = len(self._keep)
        n = np.array([rp.shape[0] - 1 for rp, _, _ in self._keep], dtype=np.int64)
        lp, ip, dp = C.POINTER(C.c_long), C.POINTER(C.c_int), C.POINTER(C.c_double)
        rps = (lp * nl)(*[a[0].ctypes.data_as(lp) for a in self._keep])
        cis = (ip * nl)(*[a[1].ctypes.data_as(ip) for a in self._keep])
        vs = (dp * nl)(*[a[2].ctypes.data_as(dp) for a in self._keep])
        cc = np.ascontiguousarray(coarse_cells, np.int32)
        self.n = int(n[-1])
        self.ptr = lib().cb_mg_create(d, k, cc.ctypes.data_as(ip), nl, n.ctypes.data_as(lp), rps, cis, vs, degree,
                                      runs, ratio, safety, power_its)
        if not self.ptr:
            raise ValueError("CPU baseline hierarchy: level sizes do not match the grids")

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().cb_mg_free(self.ptr)
            self.ptr = None

    def prolongation(self, l):
        nnz = lib().cb_mg_prol_nnz(self.ptr, l)
        nf = self._keep[l + 1][0].shape[0] - 1
        rp = np.zeros(nf + 1, np.int64)
        ci = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        lp, ip, dp = C.POINTER(C.c_long), C.POINTER(C.c_int), C.POINTER(C.c_double)
        lib().cb_mg_prol_csr(self.ptr, l, rp.ctypes.data_as(lp), ci.ctypes.data_as(ip), v.ctypes.data_as(dp))
        return rp, ci, v

    def lmax(self, l):
        return lib().cb_mg_lmax(self.ptr, l)

    def vcycle(self, x, b):
        x = np.ascontiguousarray(x, np.float64).copy()
        dp = C.POINTER(C.c_double)
        lib().cb_mg_vcycle(self.ptr, x.ctypes.data_as(dp), np.ascontiguousarray(b, np.float64).ctypes.data_as(dp))
        return x

    def pcg(self, b, tol, maxit=10000, x0=None):
        """solvePCG from x0 (default 0): (x, iterations, seconds, final residual norm)"""
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        s, r = C.c_double(), C.c_double()
        dp = C.POINTER(C.c_double)
        it = lib().cb_mg_pcg(self.ptr, b.ctypes.data_as(dp), x.ctypes.data_as(dp), tol, maxit, C.byref(s),
                             C.byref(r))
        return x, it, s.value, r.value

    def spmv_seconds(self, reps=10):
        return lib().cb_mg_spmv_seconds(self.ptr, reps)
