"""Host-side mirror of the reference's Java API for the Qk solve path.

Class and method names follow the reference (ContinuousTPFESpace,
TPCellIntegral, SparseMatrix, DenseVector, IterativeSolver, the multigrid
preconditioner and its smoothers); every operation runs in libjfem_hip.so on
the GPU through the C ABI.  Reference paths are relative to the JFEM
repository, M/ = Core/src/main/java/.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Optional, Sequence

import numpy as np

from . import _native as N
from ._native import check, dptr, f32, f64, fptr, i32, iptr, lptr

# --------------------------------------------------------------------------
# context
# --------------------------------------------------------------------------
_ctx = None


class Context:
    """One HIP device and stream (jfem_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(N.lib().jfem_ctx_create(device, C.byref(h)))
        self.h = h

    def sync(self):
        check(N.lib().jfem_ctx_sync(self.h))

    def timer_start(self):
        check(N.lib().jfem_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = C.c_double()
        check(N.lib().jfem_timer_stop(self.h, C.byref(ms)))
        return ms.value


def context() -> Context:
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


# --------------------------------------------------------------------------
# quadrature rules and integrals
# --------------------------------------------------------------------------
class QuadratureRule1D:
    """QuadratureRule1D (M/tensorproduct/QuadratureRule1D.java:9-64) by id."""
    Gauss1, Gauss2, Gauss3, Gauss4, Gauss5 = 1, 2, 3, 4, 5
    GaussLobatto2, GaussLobatto3, GaussLobatto4, GaussLobatto5, GaussLobatto6 = 12, 13, 14, 15, 16


class TPCellIntegral:
    """TPCellIntegral (M/tensorproduct/TPCellIntegral.java:17-76) with a constant
    weight; default quadrature Gauss5 as in the reference (:36-39, :58-61)."""
    GRAD_GRAD = "GradGrad"
    VALUE_VALUE = "ValueValue"
    H1 = "H1"
    GRAD_VALUE = "GradValue"
    VALUE_GRAD = "ValueGrad"
    _KIND = {GRAD_GRAD: 0, VALUE_VALUE: 1, H1: 2, GRAD_VALUE: 3, VALUE_GRAD: 4}
    _VECTOR = (GRAD_VALUE, VALUE_GRAD)

    def __init__(self, *args, quadratureRule1D: int = QuadratureRule1D.Gauss5):
        # TPCellIntegral(name) | TPCellIntegral(weight, name) | (weight, name, rule)
        if len(args) == 1:
            weight, name = 1.0, args[0]
        elif len(args) == 2:
            weight, name = args
        elif len(args) == 3:
            weight, name, quadratureRule1D = args
        else:
            raise TypeError("TPCellIntegral(name) or TPCellIntegral(weight, name[, rule])")
        if name not in self._KIND:
            raise N.UnsupportedOperationException(3, "unknown integral name")
        if name in self._VECTOR:
            # TPCellIntegral ctor (:50-55): a Vector weight (constant here)
            if isinstance(weight, (int, float)) or np.ndim(weight) != 1:
                raise N.IllegalArgumentException(1, f"{name} needs a Vector weight")
            self.weight = [float(w) for w in weight]
        elif not isinstance(weight, (int, float)):
            raise N.IllegalArgumentException(1, f"{name} needs a Double weight")
        else:
            self.weight = float(weight)
        self.name = name
        self.quadratureRule1D = quadratureRule1D

    @property
    def kind(self):
        return self._KIND[self.name]


class SinusoidSource:
    """f(x) = sum_j amp[j] prod_a sin(kw[j][a] pi x_a)."""

    def __init__(self, kw, amp):
        self.kw = i32(kw)
        self.amp = f64(amp)
        if self.kw.ndim != 2 or self.kw.shape[0] != self.amp.shape[0]:
            raise ValueError("kw must be (nterms, dim) and amp (nterms,)")

    def __call__(self, x):
        x = np.asarray(x, float)
        s = 0.0
        for j in range(len(self.amp)):
            t = self.amp[j]
            for a in range(x.shape[-1]):
                t = t * np.sin(self.kw[j, a] * np.pi * x[..., a])
            s = s + t
        return s


class TPRightHandSideIntegral:
    """TPRightHandSideIntegral VALUE (M/tensorproduct/TPRightHandSideIntegral.java:22-31)."""
    VALUE = "Value"

    def __init__(self, rightHandSide: SinusoidSource, name: str = "Value",
                 quadratureRule1D: int = QuadratureRule1D.Gauss5):
        if name != self.VALUE:
            raise N.UnsupportedOperationException(3, "unknown rhs integral")
        if not isinstance(rightHandSide, SinusoidSource):
            raise N.UnsupportedOperationException(3, "GPU right-hand sides are sinusoid sums")
        self.rightHandSide = rightHandSide
        self.quadratureRule1D = quadratureRule1D


# --------------------------------------------------------------------------
# vectors and matrices
# --------------------------------------------------------------------------
class DenseVector:
    """Device-resident FP64 vector (DenseVector, M/linalg/DenseVector.java)."""

    def __init__(self, n_or_values, ctx: Optional[Context] = None):
        self.ctx = ctx or context()
        if np.isscalar(n_or_values):
            n, values = int(n_or_values), None
        else:
            values = f64(n_or_values)
            n = values.shape[0]
        h = C.c_void_p()
        check(N.lib().jfem_vec_create(self.ctx.h, n, C.byref(h)))
        self.h = h
        self.n = n
        if values is not None and n:
            check(N.lib().jfem_vec_upload(h, dptr(values)))

    def __del__(self):
        h = getattr(self, "h", None)
        if h and N._lib is not None:
            N.lib().jfem_vec_destroy(h)
            self.h = None

    def getLength(self):
        return self.n

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n)
        if self.n:
            check(N.lib().jfem_vec_download(self.h, dptr(out)))
        return out

    def set_values(self, values):
        v = f64(values)
        if v.shape[0] != self.n:
            raise N.IllegalArgumentException(1, "Vectors have different size")
        check(N.lib().jfem_vec_upload(self.h, dptr(v)))

    def fill(self, value: float):
        check(N.lib().jfem_vec_fill(self.h, float(value)))

    def inner(self, other: "DenseVector") -> float:
        r = C.c_double()
        check(N.lib().jfem_vec_dot(self.h, other.h, C.byref(r)))
        return r.value

    def euclidianNorm(self) -> float:
        return float(np.sqrt(self.inner(self)))

    def copy(self) -> "DenseVector":
        v = DenseVector(self.n, self.ctx)
        v.copy_from(self)
        return v

    def copy_from(self, other: "DenseVector"):
        """self = other (device copy)"""
        check(N.lib().jfem_vec_copy(other.h, self.h))
        return self


class SparseMatrix:
    """Device CSR matrix behind the SparseMatrix / VectorMultiplyable interface."""

    def __init__(self, handle, ctx: Context, owned: bool = True, keepalive=None):
        self.h = handle
        self.ctx = ctx
        self.owned = owned
        self._keepalive = keepalive

    @classmethod
    def from_csr(cls, rowptr, cols, vals, shape=None, ctx: Optional[Context] = None):
        ctx = ctx or context()
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        if rp.ndim != 1 or rp.shape[0] < 1:
            raise N.IllegalArgumentException(1, "row pointer must be a non-empty 1-D array")
        n = rp.shape[0] - 1
        ncols = n if shape is None else shape[1]
        cols, vals = i32(cols), f64(vals)
        # the library reads rp[-1] entries of each array: refuse shorter ones here
        if cols.shape[0] < rp[-1] or vals.shape[0] < rp[-1]:
            raise N.IllegalArgumentException(1, f"CSR arrays hold fewer than rowptr[-1] = {rp[-1]} entries")
        h = C.c_void_p()
        check(N.lib().jfem_mat_from_csr(ctx.h, n, ncols, lptr(rp), iptr(cols), dptr(vals), C.byref(h)))
        return cls(h, ctx)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and self.owned and N._lib is not None:
            N.lib().jfem_mat_destroy(h)
            self.h = None

    def info(self):
        r, c, z = C.c_long(), C.c_long(), C.c_long()
        check(N.lib().jfem_mat_info(self.h, C.byref(r), C.byref(c), C.byref(z)))
        return r.value, c.value, z.value

    def getRows(self):
        return self.info()[0]

    def getCols(self):
        return self.info()[1]

    def getVectorSize(self):
        return self.getCols()

    def getTVectorSize(self):
        return self.getRows()

    def nnz(self):
        return self.info()[2]

    def to_csr(self):
        n, _, nnz = self.info()
        rp = np.zeros(n + 1, dtype=np.int64)
        cols = np.zeros(max(nnz, 1), dtype=np.int32)
        vals = np.zeros(max(nnz, 1))
        check(N.lib().jfem_mat_export_csr(self.h, lptr(rp), iptr(cols), dptr(vals)))
        return rp, cols[:nnz], vals[:nnz]

    def getCoordinateEntryList(self) -> dict:
        """SparseMatrix.getCoordinateEntryList (M/linalg/SparseMatrix.java:402-418)."""
        rp, cols, vals = self.to_csr()
        out = {}
        for r in range(len(rp) - 1):
            for p in range(rp[r], rp[r + 1]):
                out[(r, int(cols[p]))] = float(vals[p])
        return out

    def mvMul(self, v: DenseVector, out: Optional[DenseVector] = None) -> DenseVector:
        """SparseMatrix.mvMul (M/linalg/SparseMatrix.java:479-489)."""
        y = out if out is not None else DenseVector(self.getRows(), self.ctx)
        check(N.lib().jfem_spmv(self.h, v.h, y.h))
        return y

    def tvMul(self, v: DenseVector, out: Optional[DenseVector] = None) -> DenseVector:
        """SparseMatrix.tvMul (M/linalg/SparseMatrix.java:491-501): A^T v."""
        y = out if out is not None else DenseVector(self.getCols(), self.ctx)
        check(N.lib().jfem_spmv_t(self.h, v.h, y.h))
        return y

    def spmv_time_ms(self, x: DenseVector, y: DenseVector, reps: int = 20, fp32: bool = False) -> float:
        ms = C.c_double()
        check(N.lib().jfem_spmv_bench(self.h, x.h, y.h, reps, 1 if fp32 else 0, C.byref(ms)))
        return ms.value


# --------------------------------------------------------------------------
# space
# --------------------------------------------------------------------------
class ContinuousTPFESpace:
    """ContinuousTPFESpace (M/tensorproduct/ContinuousTPFESpace.java:13-50) on the GPU."""

    def __init__(self, startCoordinates: Sequence[float], endCoordinates: Sequence[float],
                 cellsPerDimension: Sequence[int], ctx: Optional[Context] = None):
        if len(startCoordinates) != len(endCoordinates) or len(startCoordinates) != len(cellsPerDimension):
            raise N.IllegalArgumentException(1, "Dimensions of coordinates or cell counts do not match")
        self.ctx = ctx or context()
        self.start = f64(startCoordinates)
        self.end = f64(endCoordinates)
        self.cells = i32(cellsPerDimension)
        self.dim = len(self.cells)
        self.h = None
        self.degree = None
        self.systemMatrix: Optional[SparseMatrix] = None
        self.rhs: Optional[DenseVector] = None
        self.fixedNodes: set = set()

    def __del__(self):
        if getattr(self, "h", None) and N._lib is not None:
            N.lib().jfem_space_destroy(self.h)
            self.h = None

    def assembleCells(self):
        pass

    def assembleFunctions(self, polynomialDegree: int):
        h = C.c_void_p()
        check(N.lib().jfem_space_create(self.ctx.h, self.dim, iptr(self.cells), dptr(self.start), dptr(self.end),
                                        polynomialDegree, C.byref(h)))
        self.h = h
        self.degree = polynomialDegree

    def _need(self):
        if self.h is None:
            raise N.IllegalStateException(2, "assembleFunctions has not been called")

    def getDimension(self):
        return self.dim

    def ndofs(self) -> int:
        self._need()
        return N.lib().jfem_space_ndofs(self.h)

    def cell_dofs(self) -> np.ndarray:
        self._need()
        nc = int(np.prod(self.cells.astype(np.int64)))
        out = np.zeros(nc * (self.degree + 1) ** self.dim, dtype=np.int32)
        check(N.lib().jfem_space_cell_dofs(self.h, iptr(out)))
        return out.reshape(nc, -1)

    def boundary_nodes(self) -> np.ndarray:
        self._need()
        out = np.zeros(self.ndofs(), dtype=np.int32)
        cnt = C.c_long()
        check(N.lib().jfem_space_boundary_nodes(self.h, iptr(out), C.byref(cnt)))
        return out[:cnt.value].copy()

    def node_coords(self) -> np.ndarray:
        self._need()
        out = np.zeros(self.ndofs() * self.dim)
        check(N.lib().jfem_space_node_coords(self.h, dptr(out)))
        return out.reshape(-1, self.dim)

    def initializeSystemMatrix(self):
        self._need()
        self.systemMatrix = None

    def initializeRhs(self):
        self._need()
        self.rhs = DenseVector(self.ndofs(), self.ctx)

    def getSystemMatrix(self) -> SparseMatrix:
        return self.systemMatrix

    def getRhs(self) -> DenseVector:
        return self.rhs

    def getFixedNodeIndices(self) -> set:
        return self.fixedNodes

    def assemble_matrix(self, cellIntegrals: Sequence[TPCellIntegral], cell_coeff=None) -> SparseMatrix:
        self._need()
        if not cellIntegrals:
            raise N.IllegalArgumentException(1, "no cell integrals")
        rules = {ci.quadratureRule1D for ci in cellIntegrals}
        if len(rules) != 1:
            raise N.UnsupportedOperationException(3, "all integrals must share one quadrature rule")
        kinds = i32([ci.kind for ci in cellIntegrals])
        packed = []
        for ci in cellIntegrals:
            if isinstance(ci.weight, list):
                if len(ci.weight) != self.dim:
                    raise N.IllegalArgumentException(1, f"{ci.name} weight must have {self.dim} components")
                packed += ci.weight
            else:
                packed.append(ci.weight)
        w = f64(packed)
        cc = _cell_array(cell_coeff, self, "cell_coeff")
        h = C.c_void_p()
        check(N.lib().jfem_assemble(self.h, len(kinds), iptr(kinds), dptr(w), dptr(cc) if cc is not None else None,
                                    rules.pop(), C.byref(h)))
        return SparseMatrix(h, self.ctx)

    def writeCellIntegralsToRhs(self, rhsIntegrals: Sequence[TPRightHandSideIntegral], d: DenseVector):
        for ri in rhsIntegrals:
            src = ri.rightHandSide
            tmp = DenseVector(self.ndofs(), self.ctx)
            check(N.lib().jfem_assemble_rhs_sines(self.h, len(src.amp), iptr(src.kw), dptr(src.amp),
                                                  ri.quadratureRule1D, tmp.h))
            check(N.lib().jfem_vec_axpby(1.0, tmp.h, 1.0, d.h))

    def evaluateCellIntegrals(self, cellIntegrals, rightHandSideIntegrals, cell_coeff=None):
        """MatrixFESpace.evaluateCellIntegrals (M/basic/MatrixFESpace.java:22-27)."""
        if self.rhs is None:
            self.initializeRhs()
        self.systemMatrix = self.assemble_matrix(cellIntegrals, cell_coeff)
        if rightHandSideIntegrals:
            self.writeCellIntegralsToRhs(rightHandSideIntegrals, self.rhs)

    def setBoundaryValues(self, boundaryValues: Optional[Callable] = None):
        """AcceptsBoundaryValues.setBoundaryValues (M/basic/AcceptsBoundaryValues.java:20-51)
        over the whole boundary; boundaryValues(x) -> float, None = zero data."""
        nodes = self.boundary_nodes()
        vals = None
        if boundaryValues is not None:
            xyz = self.node_coords()
            vals = f64([boundaryValues(xyz[i]) for i in nodes])
        check(N.lib().jfem_apply_dirichlet(self.systemMatrix.h, self.rhs.h, len(nodes), iptr(nodes),
                                           dptr(vals) if vals is not None else None))
        self.fixedNodes.update(int(i) for i in nodes)


class ScalarFESpaceFunction:
    """ScalarFESpaceFunction(functions, coefficients) (M/basic/ScalarFESpaceFunction.java:8-52):
    the FE function sum_i c_i phi_i of a ContinuousTPFESpace, evaluated on the GPU."""

    def __init__(self, space: ContinuousTPFESpace, coefficients: DenseVector):
        space._need()
        if coefficients.n != space.ndofs():
            raise N.IllegalArgumentException(1, "coefficient vector size differs from the space")
        self.space, self.coefficients = space, coefficients

    def getDomainDimension(self) -> int:
        return self.space.dim

    def value(self, points) -> np.ndarray:
        """values at an (npts, dim) array of points (or one point)"""
        pts = f64(points).reshape(-1, self.space.dim)
        out = np.zeros(pts.shape[0])
        check(N.lib().jfem_space_evaluate(self.space.h, self.coefficients.h, pts.shape[0], dptr(pts), dptr(out)))
        return out


def normL2Difference(f1: ScalarFESpaceFunction, f2: ScalarFESpaceFunction, points) -> float:
    """ConvergenceOrderEstimator.normL2Difference (Core/src/test/java/examples/ConvergenceOrderEstimator.java:49-57):
    sqrt(sum over the points of (f1(p) - f2(p))^2)"""
    if f1.space.dim != f2.space.dim:
        raise N.IllegalArgumentException(1, "the two functions have different dimensions")
    pts = f64(points).reshape(-1, f1.space.dim)
    out = C.c_double()
    check(N.lib().jfem_fe_l2_difference(f1.space.h, f1.coefficients.h, f2.space.h, f2.coefficients.h, pts.shape[0],
                                        dptr(pts), C.byref(out)))
    return out.value


# --------------------------------------------------------------------------
# multigrid
# --------------------------------------------------------------------------
def _smooth(kind, operator, rhs, iterate, runs, omega, degree=0, lmin=0.0, lmax=0.0):
    check(N.lib().jfem_smooth(operator.h, kind, runs, omega, degree, lmin, lmax, rhs.h, iterate.h))
    return iterate


class RichardsonSmoother:
    """RichardsonSmoother(omega, runs) (M/multigrid/RichardsonSmoother.java:13-50)."""
    kind = 0

    def __init__(self, omega: float, runs: int):
        self.omega, self.runs = omega, runs

    def smooth(self, operator: "SparseMatrix", rhs: "DenseVector", iterate: "DenseVector") -> "DenseVector":
        """Smoother.smooth on any matrix (iterate updated in place and returned)"""
        return _smooth(self.kind, operator, rhs, iterate, self.runs, self.omega)


class JacobiSmoother:
    """JacobiSmoother(runs, omega) (M/multigrid/JacobiSmoother.java:12-50)."""
    kind = 1

    def __init__(self, runs: int, omega: float):
        self.runs, self.omega = runs, omega

    def smooth(self, operator: "SparseMatrix", rhs: "DenseVector", iterate: "DenseVector") -> "DenseVector":
        """Smoother.smooth on any matrix (iterate updated in place and returned)"""
        return _smooth(self.kind, operator, rhs, iterate, self.runs, self.omega)


class ChebyshevSmoother:
    """Chebyshev-Jacobi smoother of the given degree, applied `runs` times."""
    kind = 2

    def __init__(self, degree: int = 4, runs: int = 2, ratio: float = 20.0, safety: float = 1.1, power_its: int = 30):
        self.degree, self.runs, self.ratio, self.safety, self.power_its = degree, runs, ratio, safety, power_its
        self.omega = 1.0

    def smooth(self, operator: "SparseMatrix", rhs: "DenseVector", iterate: "DenseVector",
               lmax: float = None) -> "DenseVector":
        """Chebyshev-Jacobi on [lmax / ratio, lmax] of D^-1 A (lmax given by the caller)"""
        if lmax is None:
            raise N.IllegalArgumentException(1, "Chebyshev smoothing needs the upper bound lmax of D^-1 A")
        return _smooth(self.kind, operator, rhs, iterate, self.runs, 1.0, self.degree, lmax / self.ratio, lmax)


FP64, FP32 = 0, 1


def _cell_array(values, space, what):
    """float64 copy of a per-cell array, checked against the space's cell count
    (the C ABI reads exactly that many entries)"""
    if values is None:
        return None
    a = f64(values).reshape(-1)
    nc = int(np.prod(space.cells.astype(np.int64)))
    if a.shape[0] != nc:
        raise N.IllegalArgumentException(1, f"{what} has {a.shape[0]} entries, the finest grid has {nc} cells")
    return a


class MGPreconditioner:
    """Geometric multigrid preconditioner over nested ContinuousTPFESpaces
    (MGPreconditionerSpace + MGPreconditionerInterface.mvMul,
    M/multigrid/MGPreconditionerInterface.java:352-379).  Levels: coarsest first."""

    def __init__(self, spaces: Sequence[ContinuousTPFESpace], cellIntegrals: Sequence[TPCellIntegral], smoother,
                 cycles: int = 3, fmg=True, precision: int = FP64, cell_coeff=None, galerkin: bool = False,
                 raw_transfers: bool = False, matrix_free="auto"):
        self.ctx = spaces[0].ctx
        self.spaces = list(spaces)
        opts = N.MgOpts()
        opts.smoother = smoother.kind
        opts.runs = smoother.runs
        opts.omega = getattr(smoother, "omega", 1.0)
        opts.degree = getattr(smoother, "degree", 0)
        opts.cheb_ratio = getattr(smoother, "ratio", 0.0)
        opts.cheb_safety = getattr(smoother, "safety", 1.1)
        opts.power_its = getattr(smoother, "power_its", 30)
        opts.cycles = cycles
        opts.fmg = 2 if fmg == 2 else (1 if fmg else 0)
        opts.raw_transfers = 1 if raw_transfers else 0
        opts.precision = precision
        opts.galerkin = 1 if galerkin else 0   # coarse systems P^T A P (AMGFromMatrix) instead of re-assembly
        # level operators by sum factorisation in the cycles: True everywhere it applies, "auto" on
        # 3D grids only, False never
        opts.matrix_free = 2 if matrix_free == "auto" else (1 if matrix_free else 0)
        if any(isinstance(ci.weight, list) for ci in cellIntegrals):
            raise N.UnsupportedOperationException(3, "multigrid hierarchies take scalar-weighted integrals")
        kinds = i32([ci.kind for ci in cellIntegrals])
        w = f64([ci.weight for ci in cellIntegrals])
        arr = (C.c_void_p * len(spaces))(*[s.h for s in spaces])
        cc = _cell_array(cell_coeff, spaces[-1], "cell_coeff")
        h = C.c_void_p()
        check(N.lib().jfem_mg_create(arr, len(spaces), len(kinds), iptr(kinds), dptr(w),
                                     dptr(cc) if cc is not None else None, cellIntegrals[0].quadratureRule1D,
                                     C.byref(opts), C.byref(h)))
        self.h = h
        self.nlev = len(spaces)

    def __del__(self):
        if getattr(self, "h", None) and N._lib is not None:
            N.lib().jfem_mg_destroy(self.h)
            self.h = None

    def getSystem(self, level: int) -> SparseMatrix:
        h = C.c_void_p()
        check(N.lib().jfem_mg_level_matrix(self.h, level, C.byref(h)))
        return SparseMatrix(h, self.ctx, owned=False, keepalive=self)

    def applySystem(self, level: int, x: DenseVector, out: Optional[DenseVector] = None) -> DenseVector:
        """getSystem(level).mvMul(x) with the operator the cycles use (matrix-free where enabled)"""
        y = out if out is not None else DenseVector(x.n, self.ctx)
        check(N.lib().jfem_mg_level_apply(self.h, level, x.h, y.h))
        return y

    def isMatrixFree(self, level: int) -> bool:
        v = C.c_int()
        check(N.lib().jfem_mg_level_matrix_free(self.h, level, C.byref(v)))
        return bool(v.value)

    def level_step_time_ms(self, level: int, x: DenseVector, y: DenseVector, reps: int = 20) -> float:
        """mean kernel time (ms) of one Chebyshev step of `level` from x into y (FP64 hierarchies)"""
        ms = C.c_double()
        check(N.lib().jfem_mg_level_bench(self.h, level, x.h, y.h, reps, C.byref(ms)))
        return ms.value

    def getFinestSystem(self) -> SparseMatrix:
        return self.getSystem(self.nlev - 1)

    def getVectorSize(self):
        return self.spaces[-1].ndofs()

    def mvMul(self, r: DenseVector, out: Optional[DenseVector] = None) -> DenseVector:
        z = out if out is not None else DenseVector(r.n, self.ctx)
        check(N.lib().jfem_mg_apply(self.h, r.h, z.h))
        return z

    def vCycle(self, x: DenseVector, b: DenseVector) -> DenseVector:
        check(N.lib().jfem_mg_vcycle(self.h, x.h, b.h))
        return x

    def wCycle(self, x: DenseVector, b: DenseVector) -> DenseVector:
        """MGPreconditionerInterface.wCycle (:300-305): mgStepW on the finest level, x in place"""
        check(N.lib().jfem_mg_wcycle(self.h, x.h, b.h))
        return x

    STEP_PLAIN, STEP_MIN_RESIDUAL, STEP_MIN_ENERGY = 0, 1, 2

    def set_learned(self, net, cell_log_coeff=None, runs: int = 2, fallback_ratio: float = 0.5,
                    step_rule: int = 0):
        """attach a SmootherNet as the finest-level smoother (None detaches); step_rule 0 plain
        (default), 1 minimal residual, 2 minimal energy along the network's direction"""
        cc = _cell_array(cell_log_coeff, self.spaces[-1], "cell_log_coeff")
        self._net = net
        check(N.lib().jfem_mg_set_learned(self.h, net.h if net is not None else None,
                                          dptr(cc) if cc is not None else None, runs, fallback_ratio, step_rule))

    def set_learned_active(self, active: bool):
        check(N.lib().jfem_mg_set_learned_active(self.h, 1 if active else 0))

    LEARNED_PRE, LEARNED_POST, LEARNED_BOTH = 1, 2, 3

    def set_learned_sides(self, sides: int):
        """smoothing calls on the finest level that use the network (1 pre, 2 post, 3 both); the
        other side keeps the hierarchy's smoother"""
        check(N.lib().jfem_mg_set_learned_sides(self.h, sides))

    def learned_status(self):
        a, f = C.c_int(), C.c_int()
        check(N.lib().jfem_mg_learned_status(self.h, C.byref(a), C.byref(f)))
        return bool(a.value), bool(f.value)

    def train(self, rhs_per_field: int = 10, cycles_per_rhs: int = 10, lr: float = 1e-3, seed: int = 0, mode: int = 0):
        """train the attached network on residuals of this hierarchy's plain V-cycles
        (mode 0), on A e for white-noise errors e (mode 1), alternating (mode 2),
        on residuals at cycle entry and before post-smoothing (mode 3), or as mode 3
        with known errors and the two-grid loss |T(e - N(r))|_A^2 (mode 4; the
        history then holds the per-sample ratio |T(e - N(r))|_A^2 / |e|_A^2), or as
        mode 4 with every second right-hand side advanced by the learned V-cycle (mode 5)"""
        o = N.TrainOpts(rhs_per_field, cycles_per_rhs, lr, seed, mode)
        hist = np.zeros(rhs_per_field * cycles_per_rhs)
        ns = C.c_int()
        check(N.lib().jfem_nn_train_mg(self.h, C.byref(o), dptr(hist), C.byref(ns)))
        return hist[:ns.value]

    def solve(self, b: DenseVector, x: DenseVector, tol: float, maxcycles: int = 200):
        """stationary V-cycle iteration to ||b - A x|| <= tol; returns (info, residual history)"""
        info = N.SolveInfo()
        hist = np.zeros(maxcycles)
        check(N.lib().jfem_mg_solve(self.h, b.h, x.h, tol, maxcycles, C.byref(info), dptr(hist)))
        return info, hist[:info.iterations].copy()

    def lmax(self) -> np.ndarray:
        out = np.zeros(self.nlev)
        check(N.lib().jfem_mg_lmax(self.h, dptr(out)))
        return out

    def tvMul(self, v):
        raise N.UnsupportedOperationException(3, "not implemented yet")

    def prolongate(self, level: int, coarse: DenseVector) -> DenseVector:
        """prolongationOperators.get(level).mvMul(coarse): level -> level + 1"""
        out = DenseVector(self.getSystem(level + 1).info()[0], self.ctx)
        check(N.lib().jfem_mg_prolongate(self.h, level, coarse.h, out.h))
        return out

    def restrict(self, level: int, fine: DenseVector) -> DenseVector:
        """prolongationOperators.get(level).tvMul(fine): level + 1 -> level"""
        out = DenseVector(self.getSystem(level).info()[0], self.ctx)
        check(N.lib().jfem_mg_restrict(self.h, level, fine.h, out.h))
        return out


# --------------------------------------------------------------------------
# iterative solvers
# --------------------------------------------------------------------------
def MGSpacePreconditioner(spaces, cellIntegrals, smoother, precision: int = FP64, cell_coeff=None):
    """MGSpace + MGPReconditioner2 (M/multigrid/MGSpace.java:25-206, MGPReconditioner2.java:38-76):
    the older multigrid API, transfers by the raw prolongation matrix and the application
    v + fullVCycleSolver(v - A v)"""
    return MGPreconditioner(spaces, cellIntegrals, smoother, cycles=1, fmg=2, precision=precision,
                            cell_coeff=cell_coeff, raw_transfers=True)


class IterativeSolver:
    """IterativeSolver (M/linalg/IterativeSolver.java:11-402): stateful, keeps
    `iterations` and `lastSolution` (next start iterate) like the reference."""

    def __init__(self, showProgress: bool = False):
        self.iterations = 0
        self._last: Optional[DenseVector] = None    # device copy of the last solution
        self.info = N.SolveInfo()
        self.residual_history: Optional[np.ndarray] = None

    @property
    def lastSolution(self) -> Optional[np.ndarray]:
        """the last solution (IterativeSolver.lastSolution, :17,98), the next start iterate;
        kept on the device and downloaded only when read here"""
        return None if self._last is None else self._last.numpy()

    @lastSolution.setter
    def lastSolution(self, values):
        self._last = None if values is None else DenseVector(np.asarray(values, float))

    def _start(self, n, ctx, default_e0=True):
        if self._last is not None:
            if self._last.n != n:
                raise N.IllegalArgumentException(1, "Incompatible sizes")
            return self._last.copy()
        x0 = np.zeros(n)
        if default_e0 and n:
            x0[0] = 1.0
        return DenseVector(x0, ctx)

    def _finish(self, x: DenseVector):
        self.iterations = self.info.iterations
        if self._last is None or self._last.n != x.n or self._last.ctx is not x.ctx:
            self._last = DenseVector(x.n, x.ctx)
        self._last.copy_from(x)
        return x

    def solveCG(self, operator: SparseMatrix, rhs: DenseVector, tol: float, maxit: int = 2 ** 31 - 1,
                x0: Optional[DenseVector] = None) -> DenseVector:
        x = x0 if x0 is not None else self._start(rhs.n, operator.ctx)
        hist = np.zeros(min(rhs.n, maxit) + 1)
        check(N.lib().jfem_solve_cg(operator.h, None, rhs.h, x.h, tol, maxit, C.byref(self.info), dptr(hist)))
        self.residual_history = hist[:self.info.iterations].copy()
        return self._finish(x)

    def solvePCG(self, operator: SparseMatrix, preconditioner: MGPreconditioner, rhs: DenseVector, tol: float,
                 maxit: int = 2 ** 31 - 1, x0: Optional[DenseVector] = None) -> DenseVector:
        x = x0 if x0 is not None else self._start(rhs.n, operator.ctx)
        hist = np.zeros(min(rhs.n, maxit) + 1)
        check(N.lib().jfem_solve_cg(operator.h, preconditioner.h, rhs.h, x.h, tol, maxit, C.byref(self.info),
                                    dptr(hist)))
        self.residual_history = hist[:self.info.iterations].copy()
        return self._finish(x)

    def solveFCG(self, operator: SparseMatrix, preconditioner: MGPreconditioner, rhs: DenseVector, tol: float,
                 maxit: int = 100000, x0: Optional[DenseVector] = None) -> DenseVector:
        """Flexible CG (FP64) around a multigrid preconditioner of any precision (not in the reference)."""
        x = x0 if x0 is not None else DenseVector(rhs.n, operator.ctx)
        hist = np.zeros(min(rhs.n, maxit) + 1)
        check(N.lib().jfem_solve_fcg(operator.h, preconditioner.h, rhs.h, x.h, tol, maxit, C.byref(self.info),
                                     dptr(hist)))
        self.residual_history = hist[:self.info.iterations].copy()
        return self._finish(x)

    def solveBiCGStab(self, operator, rhs, tol, maxit: int = 2 ** 31 - 1, x0=None):
        x = x0 if x0 is not None else self._start(rhs.n, operator.ctx)
        check(N.lib().jfem_solve_bicgstab(operator.h, None, rhs.h, x.h, tol, maxit, C.byref(self.info)))
        return self._finish(x)

    def solvePBiCGStab(self, operator, preconditioner, rhs, tol, maxit: int = 2 ** 31 - 1, x0=None):
        x = x0 if x0 is not None else self._start(rhs.n, operator.ctx)
        check(N.lib().jfem_solve_bicgstab(operator.h, preconditioner.h, rhs.h, x.h, tol, maxit,
                                          C.byref(self.info)))
        return self._finish(x)

    def solveGMRES(self, operator, rhs, tol, restart: int = 150, max_cycles: int = 200, x0=None):
        x = x0 if x0 is not None else self._start(rhs.n, operator.ctx)
        check(N.lib().jfem_solve_gmres(operator.h, None, rhs.h, x.h, tol, restart, max_cycles, C.byref(self.info)))
        return self._finish(x)

    def solvePGMRES(self, operator, preconditioner, rhs, tol, restart: int = 150, max_cycles: int = 200, x0=None):
        x = x0 if x0 is not None else self._start(rhs.n, operator.ctx, default_e0=False)
        check(N.lib().jfem_solve_gmres(operator.h, preconditioner.h, rhs.h, x.h, tol, restart, max_cycles,
                                       C.byref(self.info)))
        return self._finish(x)


# --------------------------------------------------------------------------
# learned smoother network
# --------------------------------------------------------------------------
class GMRES2:
    """GMRES2 (M/linalg/GMRES2.java:7-104): solve(A, b) or solve(A, M, b) with a multigrid
    preconditioner M; the iterate persists between calls as in the reference (first start e_0);
    stops when the recursively updated ||r|| <= tol (absolute).  maxit bounds the iterations
    (the reference loops until convergence)."""

    def __init__(self, tol: float, show: bool = False, maxit: int = 100000):
        self.tol, self.maxit = tol, maxit
        self.iterate: Optional[DenseVector] = None
        self.info = None

    def solve(self, A: SparseMatrix, *args) -> DenseVector:
        if len(args) == 1:
            M, b = None, args[0]
        elif len(args) == 2:
            M, b = args
        else:
            raise TypeError("solve(A, b) or solve(A, M, b)")
        if self.iterate is None:
            e0 = np.zeros(b.n)
            e0[0] = 1.0
            self.iterate = DenseVector(e0, A.ctx)
        info = N.SolveInfo()
        check(N.lib().jfem_solve_gmres2(A.h, M.h if M is not None else None, b.h, self.iterate.h, self.tol,
                                        self.maxit, C.byref(info)))
        self.info = info
        return self.iterate


class SmootherNet:
    """Convolutional network cin -> 32 -> 32 -> 32 -> 1 with 3^dim taps on the node
    lattice (jfem_nn): BF16 MFMA layers, FP32 master weights and Adam."""

    def __init__(self, dim: int = 2, cin: int = 2, seed: int = 0, ctx: Optional[Context] = None, handle=None):
        self.ctx = ctx or context()
        if handle is None:
            h = C.c_void_p()
            check(N.lib().jfem_nn_create(self.ctx.h, dim, cin, seed, C.byref(h)))
            handle = h
        self.h = handle
        self.dim, self.cin = dim, cin

    @classmethod
    def load(cls, path: str, ctx: Optional[Context] = None):
        ctx = ctx or context()
        h = C.c_void_p()
        check(N.lib().jfem_nn_load(ctx.h, path.encode(), C.byref(h)))
        net = cls.__new__(cls)
        net.ctx, net.h = ctx, h
        n = N.lib().jfem_nn_num_params(h)
        # recover dim/cin from the parameter count
        for dim in (1, 2, 3):
            for cin in range(1, 9):
                T = 3 ** dim
                if T * 32 * cin + 32 + 2 * (T * 1024 + 32) + T * 32 + 1 == n:
                    net.dim, net.cin = dim, cin
        return net

    def __del__(self):
        if getattr(self, "h", None) and N._lib is not None:
            N.lib().jfem_nn_destroy(self.h)
            self.h = None

    def num_params(self) -> int:
        return N.lib().jfem_nn_num_params(self.h)

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        check(N.lib().jfem_nn_get_params(self.h, fptr(out)))
        return out

    def set_params(self, p):
        p = f32(p)
        if p.shape[0] != self.num_params():
            raise N.IllegalArgumentException(1, "wrong parameter count")
        check(N.lib().jfem_nn_set_params(self.h, fptr(p)))

    def forward(self, shape, inp) -> np.ndarray:
        shp = i32(shape)
        x = f32(inp)
        out = np.zeros(int(np.prod(shape)), dtype=np.float32)
        check(N.lib().jfem_nn_forward_host(self.h, iptr(shp), fptr(x), fptr(out)))
        return out

    def gradient(self, shape, inp, gout) -> np.ndarray:
        shp = i32(shape)
        g = np.zeros(self.num_params(), dtype=np.float32)
        check(N.lib().jfem_nn_gradient_host(self.h, iptr(shp), fptr(f32(inp)), fptr(f32(gout)), fptr(g)))
        return g

    def adam_step(self, grad=None, lr: float = 1e-3):
        check(N.lib().jfem_nn_adam_step(self.h, fptr(f32(grad)) if grad is not None else None, lr))

    def save(self, path: str):
        check(N.lib().jfem_nn_save(self.h, path.encode()))
