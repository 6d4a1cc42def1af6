"""ctypes front end of the CPU oracle (libjfem_oracle.so built from jfem_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as a product code path.  Every function here restates the reference's
Java algorithm as cited in jfem_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libjfem_oracle.so")

GRAD_GRAD, VALUE_VALUE, H1, GRAD_VALUE, VALUE_GRAD = 0, 1, 2, 3, 4
GAUSS = {1: 1, 2: 2, 3: 3, 4: 4, 5: 5}
GAUSS_LOBATTO = {2: 12, 3: 13, 4: 14, 5: 15, 6: 16}
SM_RICHARDSON, SM_JACOBI, SM_CHEBYSHEV = 0, 1, 2

_SFUN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_int, C.c_void_p)
_VFUN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double), C.c_void_p)


class _CSR(C.Structure):
    _fields_ = [("n", C.c_long), ("nnz", C.c_long), ("rowptr", C.POINTER(C.c_long)),
                ("cols", C.POINTER(C.c_int)), ("vals", C.POINTER(C.c_double))]


class _SolveOut(C.Structure):
    _fields_ = [("iterations", C.c_int), ("fell_back", C.c_int), ("restarts", C.c_int),
                ("resnorm", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        L.orc_quad.argtypes = [C.c_int, dp, dp]
        L.orc_lagrange_value.restype = C.c_double
        L.orc_lagrange_value.argtypes = [C.c_double, C.c_int, C.c_int]
        L.orc_lagrange_deriv.restype = C.c_double
        L.orc_lagrange_deriv.argtypes = [C.c_double, C.c_int, C.c_int]
        L.orc_cells1d.argtypes = [C.c_double, C.c_double, C.c_int, dp, dp]
        L.orc_dof_table.restype = C.c_long
        L.orc_dof_table.argtypes = [C.c_int, C.c_int, ip, ip]
        L.orc_cell_integral.restype = C.c_double
        L.orc_cell_integral.argtypes = [C.c_int, C.c_int, dp, dp, C.c_int, C.c_int, ip, ip,
                                        C.c_double, _SFUN, _VFUN, C.c_void_p]
        L.orc_assemble.restype = C.POINTER(_CSR)
        L.orc_assemble.argtypes = [C.c_int, C.c_int, ip, dp, dp, C.c_int, ip, dp, dp, C.c_int]
        L.orc_assemble_w.restype = C.POINTER(_CSR)
        L.orc_assemble_w.argtypes = [C.c_int, C.c_int, ip, dp, dp, C.c_int, ip, dp, dp, dp, C.c_int]
        L.orc_csr_free.argtypes = [C.POINTER(_CSR)]
        L.orc_element_matrix.argtypes = [C.c_int, C.c_int, ip, dp, dp, C.c_long, C.c_int, ip, dp, C.c_int, dp]
        L.orc_assemble_rhs_sines.argtypes = [C.c_int, C.c_int, ip, dp, dp, C.c_int, C.c_int, ip, dp, dp]
        L.orc_node_coords.argtypes = [C.c_int, C.c_int, ip, dp, dp, dp]
        L.orc_dirichlet.argtypes = [C.c_long, lp, ip, dp, dp, C.c_long, ip, dp]
        L.orc_boundary_nodes.restype = C.c_long
        L.orc_boundary_nodes.argtypes = [C.c_int, C.c_int, ip, ip]
        L.orc_spmv.argtypes = [C.c_long, lp, ip, dp, dp, dp]
        L.orc_lu_solve.argtypes = [C.c_int, dp, dp, dp]
        L.orc_prolongation.restype = C.POINTER(_CSR)
        L.orc_prolongation.argtypes = [C.c_int, C.c_int, ip, dp, dp]
        L.orc_mg_create.restype = C.c_void_p
        L.orc_mg_create.argtypes = [C.c_int, C.c_int, ip, C.c_int, dp, dp, C.c_int, ip, dp, C.c_int,
                                    C.c_int, C.c_int, C.c_double, C.c_int, dp, dp, C.c_int, C.c_int]
        L.orc_mg_free.argtypes = [C.c_void_p]
        L.orc_mg_level_n.restype = C.c_long
        L.orc_mg_level_n.argtypes = [C.c_void_p, C.c_int]
        L.orc_mg_level_nnz.restype = C.c_long
        L.orc_mg_level_nnz.argtypes = [C.c_void_p, C.c_int]
        L.orc_mg_level_csr.argtypes = [C.c_void_p, C.c_int, lp, ip, dp]
        L.orc_mg_level_dinv.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_mg_prol_nnz.restype = C.c_long
        L.orc_mg_prol_nnz.argtypes = [C.c_void_p, C.c_int]
        L.orc_mg_prol_csr.argtypes = [C.c_void_p, C.c_int, lp, ip, dp]
        L.orc_mg_apply.argtypes = [C.c_void_p, dp, dp]
        L.orc_mg_vcycle.argtypes = [C.c_void_p, dp, dp]
        L.orc_mg_wcycle.argtypes = [C.c_void_p, dp, dp]
        L.orc_mg_set_raw.argtypes = [C.c_void_p, C.c_int]
        L.orc_mg_smooth.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.orc_set_smooth_plant.argtypes = [C.c_int]
        L.orc_mg_set_interval.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.orc_solve.argtypes = [C.c_int, C.c_long, lp, ip, dp, C.c_int, C.c_void_p, lp, ip, dp,
                                dp, dp, C.c_double, C.c_long, C.c_int, C.POINTER(_SolveOut), dp]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _d(a):
    return _p(a, C.c_double)


def _i(a):
    return _p(a, C.c_int)


def _l(a):
    return _p(a, C.c_long)


def quad(rule):
    pts = np.zeros(8)
    wts = np.zeros(8)
    n = lib().orc_quad(rule, _d(pts), _d(wts))
    return pts[:n].copy(), wts[:n].copy()


def lagrange_value(pos, l, k):
    return lib().orc_lagrange_value(pos, l, k)


def lagrange_deriv(pos, l, k):
    return lib().orc_lagrange_deriv(pos, l, k)


def cells1d(start, end, n):
    s = np.zeros(n)
    e = np.zeros(n)
    lib().orc_cells1d(start, end, n, _d(s), _d(e))
    return s, e


def dof_table(cells, k):
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    d = len(cells)
    out = np.zeros(int(np.prod(cells)) * (k + 1) ** d, dtype=np.int32)
    n = lib().orc_dof_table(d, k, _i(cells), _i(out))
    return n, out.reshape(int(np.prod(cells)), (k + 1) ** d)


def cell_integral(start, end, k, kind, rule, lu, lv, wconst=1.0, wfun=None, vfun=None):
    """One TPCellIntegral value on the cell prod [start_a, end_a]; lu/lv local multi-indices."""
    s = np.ascontiguousarray(start, dtype=np.float64)
    e = np.ascontiguousarray(end, dtype=np.float64)
    d = len(s)
    lu = np.ascontiguousarray(lu, dtype=np.int32)
    lv = np.ascontiguousarray(lv, dtype=np.int32)
    sf = _SFUN(lambda x, dd, c: float(wfun([x[i] for i in range(dd)]))) if wfun else _SFUN(0)
    if vfun:
        def _vf(x, dd, out, c):
            v = vfun([x[i] for i in range(dd)])
            for i in range(dd):
                out[i] = v[i]
        vf = _VFUN(_vf)
    else:
        vf = _VFUN(0)
    return lib().orc_cell_integral(d, k, _d(s), _d(e), kind, rule, _i(lu), _i(lv), wconst, sf, vf, None)


class CSR:
    """Host CSR (int64 rowptr, int32 cols, float64 values)."""

    def __init__(self, n, rowptr, cols, vals, ncols=None):
        self.n = n
        self.ncols = n if ncols is None else ncols
        self.rowptr = rowptr
        self.cols = cols
        self.vals = vals

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.cols, self.rowptr), shape=(self.n, self.ncols))

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        lib().orc_spmv(self.n, _l(self.rowptr), _i(self.cols), _d(self.vals), _d(x), _d(y))
        return y


def _take_csr(ptr, ncols=None):
    c = ptr.contents
    n, nnz = c.n, c.nnz
    rowptr = np.ctypeslib.as_array(c.rowptr, shape=(n + 1,)).copy()
    cols = np.ctypeslib.as_array(c.cols, shape=(max(nnz, 1),))[:nnz].copy()
    vals = np.ctypeslib.as_array(c.vals, shape=(max(nnz, 1),))[:nnz].copy()
    lib().orc_csr_free(ptr)
    return CSR(n, rowptr, cols, vals, ncols)


def assemble(cells, k, start, end, kinds, weights, cell_coeff=None, rule=5):
    """weights: one float per integral, or a sequence of dim floats for GRAD_VALUE / VALUE_GRAD"""
    if any(kk in (GRAD_VALUE, VALUE_GRAD) for kk in kinds):
        d = len(cells)
        wc = np.array([0.0 if np.ndim(w) else float(w) for w in weights])
        wv = np.zeros(len(kinds) * d)
        for i, w in enumerate(weights):
            if np.ndim(w):
                wv[i * d:(i + 1) * d] = w
        cells_ = np.ascontiguousarray(cells, dtype=np.int32)
        kinds_ = np.ascontiguousarray(kinds, dtype=np.int32)
        cc = None if cell_coeff is None else np.ascontiguousarray(cell_coeff, dtype=np.float64)
        ptr = lib().orc_assemble_w(d, k, _i(cells_), _d(np.asarray(start, float)), _d(np.asarray(end, float)),
                                   len(kinds_), _i(kinds_), _d(wc), _d(wv), _d(cc) if cc is not None else None, rule)
        if not ptr:
            raise ValueError("oracle assembly failed")
        return _take_csr(ptr)
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    d = len(cells)
    s = np.ascontiguousarray(start, dtype=np.float64)
    e = np.ascontiguousarray(end, dtype=np.float64)
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    cc = None if cell_coeff is None else np.ascontiguousarray(cell_coeff, dtype=np.float64)
    ptr = lib().orc_assemble(d, k, _i(cells), _d(s), _d(e), len(kinds), _i(kinds), _d(w),
                             _d(cc) if cc is not None else None, rule)
    if not ptr:
        raise ValueError("oracle assembly failed")
    return _take_csr(ptr)


def element_matrix(cells, k, start, end, cell, kinds, weights, rule=5):
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    d = len(cells)
    nloc = (k + 1) ** d
    out = np.zeros(nloc * nloc)
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    lib().orc_element_matrix(d, k, _i(cells), _d(np.asarray(start, float)), _d(np.asarray(end, float)),
                             cell, len(kinds), _i(kinds), _d(w), rule, _d(out))
    return out.reshape(nloc, nloc)


def rhs_sines(cells, k, start, end, kw, amp, rule=5):
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    d = len(cells)
    n = int(np.prod(np.asarray(cells, np.int64) * k + 1))
    rhs = np.zeros(n)
    kw = np.ascontiguousarray(kw, dtype=np.int32)
    amp = np.ascontiguousarray(amp, dtype=np.float64)
    rc = lib().orc_assemble_rhs_sines(d, k, _i(cells), _d(np.asarray(start, float)), _d(np.asarray(end, float)),
                                      rule, len(amp), _i(kw), _d(amp), _d(rhs))
    if rc < 0:
        raise ValueError("oracle right-hand side failed (Gauss-Lobatto rules are refused)")
    return rhs


def fe_value(cells, k, start, end, coeffs, points):
    """ScalarFESpaceFunction.value (M/basic/ScalarFESpaceFunction.java:40-45) by
    brute force over all shape functions, for small grids: each function is
    evaluated in the first of its cells (cell order) that contains the point
    (ScalarShapeFunction.value, M/basic/ScalarShapeFunction.java:26-31;
    Cell1D.isInCell with 1e-8 slack), as the product of its 1D Lagrange values
    there; sum over the functions in DOF order."""
    d = len(cells)
    n, table = dof_table(cells, k)
    s1 = [cells1d(start[a], end[a], cells[a])[0] for a in range(d)]
    e1 = [cells1d(start[a], end[a], cells[a])[1] for a in range(d)]
    ncell = table.shape[0]
    # cell multi-index in IntCoordinates order (last axis fastest)
    cidx = np.zeros((ncell, d), dtype=np.int64)
    r = np.arange(ncell)
    for a in range(d - 1, -1, -1):
        cidx[:, a] = r % cells[a]
        r //= cells[a]
    fcells = [[] for _ in range(n)]   # (cell, local index) per function, in cell order
    for c in range(ncell):
        for li, f in enumerate(table[c]):
            fcells[f].append((c, li))
    pts = np.asarray(points, float).reshape(-1, d)
    out = np.zeros(len(pts))
    for i, x in enumerate(pts):
        v = 0.0
        for f in range(n):
            for c, li in fcells[f]:
                ca = cidx[c]
                if all(s1[a][ca[a]] - 1e-8 <= x[a] <= e1[a][ca[a]] + 1e-8 for a in range(d)):
                    phi = 1.0
                    rr = li
                    for a in range(d):
                        la = rr % (k + 1)
                        rr //= k + 1
                        s, e = s1[a][ca[a]], e1[a][ca[a]]
                        phi *= lagrange_value((x[a] - s) / (e - s), la, k)
                    v += coeffs[f] * phi
                    break
        out[i] = v
    return out


def node_coords(cells, k, start, end):
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    d = len(cells)
    n = int(np.prod(np.asarray(cells, np.int64) * k + 1))
    xyz = np.zeros(n * d)
    lib().orc_node_coords(d, k, _i(cells), _d(np.asarray(start, float)), _d(np.asarray(end, float)), _d(xyz))
    return xyz.reshape(n, d)


def boundary_nodes(cells, k):
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    n = int(np.prod(np.asarray(cells, np.int64) * k + 1))
    out = np.zeros(n, dtype=np.int32)
    nb = lib().orc_boundary_nodes(len(cells), k, _i(cells), _i(out))
    return out[:nb].copy()


def dirichlet(A: CSR, rhs, nodes, g=None):
    """In-place symmetric elimination (AcceptsMatrixBoundaryValues.overWriteValue)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
    rc = lib().orc_dirichlet(A.n, _l(A.rowptr), _i(A.cols), _d(A.vals), _d(rhs), len(nodes), _i(nodes),
                             _d(gg) if gg is not None else None)
    if rc:
        raise ValueError("Dirichlet node without diagonal entry")


def lu_solve(Adense, b):
    A = np.ascontiguousarray(Adense, dtype=np.float64).copy()
    n = A.shape[0]
    x = np.zeros(n)
    if lib().orc_lu_solve(n, _d(A), _d(np.ascontiguousarray(b, float)), _d(x)):
        raise ValueError("singular")
    return x


def prolongation(cells_coarse, k, start, end):
    cc = np.ascontiguousarray(cells_coarse, dtype=np.int32)
    ptr = lib().orc_prolongation(len(cc), k, _i(cc), _d(np.asarray(start, float)), _d(np.asarray(end, float)))
    ncoarse = int(np.prod(np.asarray(cc, np.int64) * k + 1))
    return _take_csr(ptr, ncoarse)


METHOD = {"cg": 0, "bicgstab": 1, "gmres": 2, "gmres2": 3}


def solve(method, A: CSR, b, x0, tol, precond=None, maxit=10 ** 9, restart=150, hist=None):
    """Reference Krylov solvers.  precond: None, 'identity', an MG object or a CSR."""
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = _SolveOut()
    pk, mgp, M = 0, None, None
    if isinstance(precond, str) and precond == "identity":
        pk = 1
    elif isinstance(precond, MG):
        pk, mgp = 2, precond.ptr
    elif isinstance(precond, CSR):
        pk, M = 3, precond
    h = None if hist is None else _d(hist)
    lib().orc_solve(METHOD[method], A.n, _l(A.rowptr), _i(A.cols), _d(A.vals), pk, mgp,
                    _l(M.rowptr) if M else None, _i(M.cols) if M else None, _d(M.vals) if M else None,
                    _d(b), _d(x), tol, maxit, restart, C.byref(out), h)
    return x, out.iterations, out.resnorm, out.fell_back


class MG:
    """Oracle geometric multigrid hierarchy (levels base_cells * 2^l)."""

    def __init__(self, base_cells, k, nlev, start, end, kinds, weights, rule=5, smoother=SM_JACOBI,
                 runs=2, omega=0.7, degree=4, lmin=None, lmax=None, cycles=1, fmg=False):
        bc = np.ascontiguousarray(base_cells, dtype=np.int32)
        kinds = np.ascontiguousarray(kinds, dtype=np.int32)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        lo = None if lmin is None else np.ascontiguousarray(lmin, dtype=np.float64)
        hi = None if lmax is None else np.ascontiguousarray(lmax, dtype=np.float64)
        self.nlev = nlev
        self.ptr = lib().orc_mg_create(len(bc), k, _i(bc), nlev, _d(np.asarray(start, float)),
                                       _d(np.asarray(end, float)), len(kinds), _i(kinds), _d(w), rule,
                                       smoother, runs, omega, degree, _d(lo) if lo is not None else None,
                                       _d(hi) if hi is not None else None, cycles,
                                       2 if (fmg == 2 and fmg is not True) else (1 if fmg else 0))

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_mg_free(self.ptr)
            self.ptr = None

    def level_n(self, l):
        return lib().orc_mg_level_n(self.ptr, l)

    def level_matrix(self, l):
        n = self.level_n(l)
        nnz = lib().orc_mg_level_nnz(self.ptr, l)
        rp = np.zeros(n + 1, dtype=np.int64)
        c = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz)
        lib().orc_mg_level_csr(self.ptr, l, _l(rp), _i(c), _d(v))
        return CSR(n, rp, c, v)

    def level_dinv(self, l):
        out = np.zeros(self.level_n(l))
        lib().orc_mg_level_dinv(self.ptr, l, _d(out))
        return out

    def prolongation(self, l):
        nf = self.level_n(l + 1)
        nnz = lib().orc_mg_prol_nnz(self.ptr, l)
        rp = np.zeros(nf + 1, dtype=np.int64)
        c = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz)
        lib().orc_mg_prol_csr(self.ptr, l, _l(rp), _i(c), _d(v))
        return CSR(nf, rp, c, v, self.level_n(l))

    def set_interval(self, l, lmin, lmax):
        lib().orc_mg_set_interval(self.ptr, l, lmin, lmax)

    def estimate_lmax(self, l, its=30):
        """power iteration on D^-1 A of level l (same recurrence as the device estimate)"""
        A = self.level_matrix(l).to_scipy()
        dinv = self.level_dinv(l)
        n = A.shape[0]
        i = np.arange(n, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = i * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x632BE59BD9B4E019)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        v = 0.5 + (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        v /= np.linalg.norm(v)
        lam = 0.0
        for _ in range(its):
            w = dinv * (A @ v)
            lam = np.linalg.norm(w)
            if lam == 0:
                break
            v = w / lam
        return lam

    def apply(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros_like(v)
        lib().orc_mg_apply(self.ptr, _d(v), _d(out))
        return out

    def vcycle(self, x, b):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().orc_mg_vcycle(self.ptr, _d(x), _d(np.ascontiguousarray(b, float)))
        return x

    def set_raw(self, raw=True):
        """MGSpace transfers: the prolongation matrix without zero boundary values"""
        lib().orc_mg_set_raw(self.ptr, 1 if raw else 0)

    def wcycle(self, x, b):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().orc_mg_wcycle(self.ptr, _d(x), _d(np.ascontiguousarray(b, float)))
        return x

    def smooth(self, l, b, x):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().orc_mg_smooth(self.ptr, l, _d(np.ascontiguousarray(b, float)), _d(x))
        return x


class planted_smoother:
    """Context manager for the closed-form smoother tests: plants a known mistake in the oracle's
    smoother (see orc_set_smooth_plant) and removes it on exit."""
    CHEB_C1, CHEB_C2, CHEB_CENTRE, JACOBI_OMEGA, RICHARDSON_EXIT = 1, 2, 3, 4, 5

    def __init__(self, plant):
        self.plant = plant

    def __enter__(self):
        lib().orc_set_smooth_plant(self.plant)
        return self

    def __exit__(self, *exc):
        lib().orc_set_smooth_plant(0)
        return False


class SplitMix64:
    """SplitMix64 stream (same constants as orc_splitmix64) with Box-Muller normals."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = seed & self.MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo, hi):
        """integer uniform in [lo, hi]"""
        return lo + self.next_u64() % (hi - lo + 1)

    def normal(self):
        import math
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
