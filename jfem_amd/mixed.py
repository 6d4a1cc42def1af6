"""Host-side mirror of the reference's Taylor-Hood (Stokes) API.

TaylorHoodSpace, the mixed cell and right-hand-side integrals, boundary data
and the pressure pin follow M/mixed/ and M/basic/AcceptsMatrixBoundaryValues;
assembly, elimination and the block-preconditioned FGMRES solve run in
libjfem_hip.so through the C ABI.  Reference paths are relative to the JFEM
repository, M/ = Core/src/main/java/.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Optional, Sequence

import numpy as np

from . import _native as N
from ._native import check, dptr, f64, i32, iptr
from .api import (ContinuousTPFESpace, Context, DenseVector, MGPreconditioner, QuadratureRule1D, SparseMatrix,
                  TPCellIntegral, context)


class TrigSource:
    """f(x) = sum_j amp[j] prod_a t_ja(kw[j][a] pi x_a) with t = sin, or cos where cos[j][a] is set."""

    def __init__(self, kw, amp, cos=None):
        self.kw = i32(kw)
        self.amp = f64(amp)
        if self.kw.ndim != 2 or self.kw.shape[0] != self.amp.shape[0]:
            raise ValueError("kw must be (nterms, dim) and amp (nterms,)")
        self.cos = i32(np.zeros_like(self.kw) if cos is None else cos)
        if self.cos.shape != self.kw.shape:
            raise ValueError("cos must have the shape of kw")

    def __call__(self, x):
        x = np.asarray(x, float)
        s = 0.0
        for j in range(len(self.amp)):
            t = self.amp[j]
            for a in range(x.shape[-1]):
                f = np.cos if self.cos[j, a] else np.sin
                t = t * f(self.kw[j, a] * np.pi * x[..., a])
            s = s + t
        return s


class VectorTrigSource:
    """A body force with one TrigSource per velocity component."""

    def __init__(self, components: Sequence[TrigSource]):
        self.components = list(components)

    def __call__(self, x):
        return np.stack([c(x) for c in self.components], axis=-1)


class TPVectorCellIntegral:
    """TPVectorCellIntegral GRAD_GRAD with a constant weight
    (M/tensorproduct/TPVectorCellIntegral.java:80-88), default quadrature Gauss5."""
    GRAD_GRAD = "GradGrad"

    def __init__(self, weight, name: str, quadratureRule1D: int = QuadratureRule1D.Gauss5):
        if name != self.GRAD_GRAD:
            raise N.UnsupportedOperationException(3, "only GRAD_GRAD velocity integrals")
        if not isinstance(weight, (int, float)):
            raise N.IllegalArgumentException(1, f"{name} needs a Double weight")
        self.weight, self.name, self.quadratureRule1D = float(weight), name, quadratureRule1D


class MixedCellIntegral:
    """MixedCellIntegral (M/mixed/MixedCellIntegral.java:11-124); fromVelocityIntegral
    wraps a velocity integral (:58-68)."""
    kind = 0

    def __init__(self, weight: float, quadratureRule1D: int):
        self.weight, self.quadratureRule1D = weight, quadratureRule1D

    @staticmethod
    def fromVelocityIntegral(vi: TPVectorCellIntegral) -> "MixedCellIntegral":
        return MixedCellIntegral(vi.weight, vi.quadratureRule1D)


class MixedTPCellIntegral:
    """MixedTPCellIntegral (M/mixed/MixedTPCellIntegral.java:17-59): DIV_VALUE weight * div(u) * p or
    VALUE_GRAD weight * u . grad(p), at (velocity, pressure) and (pressure, velocity)."""
    DIV_VALUE = "DivValue"
    VALUE_GRAD = "ValueGrad"

    def __init__(self, *args, quadratureRule1D: int = QuadratureRule1D.Gauss5):
        if len(args) == 1:
            weight, name = 1.0, args[0]
        elif len(args) == 2:
            weight, name = args
        else:
            raise TypeError("MixedTPCellIntegral(name) or MixedTPCellIntegral(weight, name)")
        if name not in (self.DIV_VALUE, self.VALUE_GRAD):
            raise N.UnsupportedOperationException(3, "Integral type not supported")
        self.weight, self.name, self.quadratureRule1D = float(weight), name, quadratureRule1D
        self.kind = 1 if name == self.DIV_VALUE else 2


class TPVectorRightHandSideIntegral:
    """TPVectorRightHandSideIntegral VALUE (M/tensorproduct/TPVectorRightHandSideIntegral.java:25-33)."""
    VALUE = "Value"

    def __init__(self, rightHandSide: VectorTrigSource, name: str = "Value",
                 quadratureRule1D: int = QuadratureRule1D.Gauss5):
        if name != self.VALUE:
            raise N.UnsupportedOperationException(3, "unknown rhs integral")
        if not isinstance(rightHandSide, VectorTrigSource):
            raise N.UnsupportedOperationException(3, "GPU body forces are sums of sine/cosine products")
        self.rightHandSide, self.quadratureRule1D = rightHandSide, quadratureRule1D


class MixedRightHandSideIntegral:
    """MixedRightHandSideIntegral.fromVelocityIntegral (M/mixed/MixedRightHandSideIntegral.java:35-74)."""

    def __init__(self, vi: TPVectorRightHandSideIntegral):
        self.velocityIntegral = vi

    @staticmethod
    def fromVelocityIntegral(vi: TPVectorRightHandSideIntegral) -> "MixedRightHandSideIntegral":
        return MixedRightHandSideIntegral(vi)


class TaylorHoodSpace:
    """TaylorHoodSpace (M/mixed/TaylorHoodSpace.java:17-92): Q(k+1)^d velocity and Qk pressure.
    Unknowns: velocity component 0, ..., d-1, then pressure, each by node point in
    CoordinateComparator order (QkQkFunction.compareTo, M/mixed/QkQkFunction.java:82-93)."""

    def __init__(self, startCoordinates: Sequence[float], endCoordinates: Sequence[float],
                 cellsPerDimension: Sequence[int], ctx: Optional[Context] = None):
        self.ctx = ctx or context()
        self.args = (list(startCoordinates), list(endCoordinates), list(cellsPerDimension))
        self.dim = len(cellsPerDimension)
        self.velocity: Optional[ContinuousTPFESpace] = None
        self.pressure: Optional[ContinuousTPFESpace] = None

    def assembleCells(self):
        pass

    def assembleFunctions(self, polynomialDegree: int):
        """pressure degree k = polynomialDegree, velocity degree k + 1 (:41-92)"""
        s, e, c = self.args
        self.velocity = ContinuousTPFESpace(s, e, c, self.ctx)
        self.velocity.assembleFunctions(polynomialDegree + 1)
        self.pressure = ContinuousTPFESpace(s, e, c, self.ctx)
        self.pressure.assembleFunctions(polynomialDegree)
        self.nv = self.velocity.ndofs()
        self.np = self.pressure.ndofs()
        self._vrank = self._lex_rank(self.velocity)
        self._prank = self._lex_rank(self.pressure)

    @staticmethod
    def _lex_rank(space: ContinuousTPFESpace) -> np.ndarray:
        out = np.zeros(space.ndofs(), dtype=np.int32)
        check(N.lib().jfem_space_lex_rank(space.h, iptr(out)))
        return out

    def _need(self):
        if self.velocity is None:
            raise N.IllegalStateException(2, "assembleFunctions has not been called")

    def getVelocitySize(self) -> int:
        self._need()
        return self.dim * self.nv

    def ndofs(self) -> int:
        self._need()
        return self.dim * self.nv + self.np

    def velocity_node_coords(self) -> np.ndarray:
        """(nv, d) node points of one velocity component in the mixed numbering's order"""
        self._need()
        xyz = np.empty((self.nv, self.dim))
        xyz[self._vrank] = self.velocity.node_coords()
        return xyz

    def pressure_node_coords(self) -> np.ndarray:
        self._need()
        xyz = np.empty((self.np, self.dim))
        xyz[self._prank] = self.pressure.node_coords()
        return xyz

    def velocity_boundary_nodes(self) -> np.ndarray:
        """node-order indices (0..nv-1) of the velocity nodes on the boundary"""
        self._need()
        return np.sort(self._vrank[self.velocity.boundary_nodes()])

    def velocity_component(self, x: DenseVector, c: int) -> np.ndarray:
        """component c of the velocity part of a mixed vector, in the scalar space's DOF order"""
        self._need()
        return x.numpy()[c * self.nv + self._vrank]

    def pressure_part(self, x: DenseVector) -> np.ndarray:
        """pressure part of a mixed vector, in the scalar pressure space's DOF order"""
        self._need()
        return x.numpy()[self.dim * self.nv + self._prank]

    def writeCellIntegralsToMatrix(self, cellIntegrals) -> SparseMatrix:
        """MatrixFESpace.writeCellIntegralsToMatrix (M/basic/MatrixFESpace.java:22-42) with
        MixedCellIntegral.evaluateCellIntegral (M/mixed/MixedCellIntegral.java:93-124)"""
        self._need()
        rules = {ci.quadratureRule1D for ci in cellIntegrals}
        if len(rules) != 1:
            raise N.UnsupportedOperationException(3, "all integrals must share one quadrature rule")
        kinds = i32([ci.kind for ci in cellIntegrals])
        w = f64([ci.weight for ci in cellIntegrals])
        h = C.c_void_p()
        check(N.lib().jfem_th_assemble(self.velocity.h, self.pressure.h, len(kinds), iptr(kinds), dptr(w),
                                       rules.pop(), C.byref(h)))
        return SparseMatrix(h, self.ctx)

    def writeCellIntegralsToRhs(self, rhsIntegrals: Sequence[MixedRightHandSideIntegral], rhs: DenseVector):
        """adds the velocity right-hand side integrals into rhs (writeCellIntegralsToRhs adds)"""
        self._need()
        for ri in rhsIntegrals:
            vi = ri.velocityIntegral
            comps = vi.rightHandSide.components
            if len(comps) != self.dim:
                raise N.IllegalArgumentException(1, "body force has the wrong number of components")
            for c, src in enumerate(comps):
                check(N.lib().jfem_th_assemble_rhs_trig(self.velocity.h, self.pressure.h, c, len(src.amp),
                                                        iptr(src.kw), iptr(src.cos), dptr(src.amp),
                                                        vi.quadratureRule1D, rhs.h))

    def writeVelocityBoundaryValuesTo(self, boundaryValues: Optional[Callable], A: SparseMatrix, rhs: DenseVector):
        """writeBoundaryValuesTo(values, f -> true, (f, sf) -> sf.hasVelocityFunction(), A, rhs)
        (M/basic/AcceptsMatrixBoundaryValues.java:62-75): every velocity node on the boundary,
        both components, eliminated with value boundaryValues(x)[c] (None = zero data)"""
        self._need()
        bn = self.velocity_boundary_nodes()
        nodes = np.concatenate([c * self.nv + bn for c in range(self.dim)]).astype(np.int32)
        vals = None
        if boundaryValues is not None:
            xyz = self.velocity_node_coords()[bn]
            g = np.array([np.asarray(boundaryValues(p), float) for p in xyz]).reshape(len(bn), self.dim)
            vals = f64(np.concatenate([g[:, c] for c in range(self.dim)]))
        check(N.lib().jfem_apply_dirichlet(A.h, rhs.h, len(nodes), iptr(nodes),
                                           dptr(vals) if vals is not None else None))

    def overWriteValue(self, index: int, value: float, A: SparseMatrix, rhs: DenseVector):
        """AcceptsMatrixBoundaryValues.overWriteValue (:25-44); creates the diagonal entry"""
        nodes = i32([index])
        check(N.lib().jfem_apply_dirichlet(A.h, rhs.h, 1, iptr(nodes), dptr(f64([value]))))


class StokesBlockPreconditioner:
    """Block lower-triangular preconditioner for the Taylor-Hood Stokes system:
    velocity block = `velocity_mg` (a multigrid hierarchy of the scalar velocity space with
    GRAD_GRAD weight `viscosity` and the same Dirichlet boundary) once per component;
    Schur complement S ~ M_p / viscosity, inverted by the lumped mass (mass_its = 0) or by
    mass_its Chebyshev steps on M_p / viscosity preconditioned by the lumped mass; the pinned
    pressure row (default: the first pressure unknown) is kept as an identity row."""

    def __init__(self, space: TaylorHoodSpace, velocity_mg: MGPreconditioner, viscosity: float,
                 pinned: Sequence[int] = (0,), mass_its: int = 0,
                 component_mgs: Sequence[MGPreconditioner] = ()):
        """component_mgs: optional copies of velocity_mg for components 1 (and 2), built on other
        Contexts; their V-cycles then run on those contexts' streams, concurrently"""
        space._need()
        if velocity_mg.getVectorSize() != space.nv:
            raise N.IllegalArgumentException(1, "velocity multigrid does not match the velocity space")
        if len(component_mgs) > space.dim - 1:
            raise N.IllegalArgumentException(1, "more component hierarchies than velocity components")
        self.space, self.mg = space, velocity_mg
        self.component_mgs = list(component_mgs)
        self.mass = space.pressure.assemble_matrix([TPCellIntegral(1.0 / viscosity, TPCellIntegral.VALUE_VALUE)])
        ones = DenseVector(np.ones(space.np), space.ctx)
        lumped_dof = self.mass.mvMul(ones).numpy()
        lumped = np.empty(space.np)
        lumped[space._prank] = lumped_dof
        scale = 1.0 / lumped
        scale[list(pinned)] = -1.0
        self.scale = DenseVector(scale, space.ctx)
        self.mass_its = mass_its
        if mass_its > 0 and space.pressure.degree != 1:
            raise N.UnsupportedOperationException(3, "mass Chebyshev bounds are known for Q1 pressure only")

    def opts(self, restart: int, maxit: int) -> N.StokesOpts:
        o = N.StokesOpts()
        o.restart, o.maxit = restart, maxit
        if self.mass_its > 0:
            # (lumped Q1 mass)^-1 Q1 mass has its spectrum in [3^-dim, 1] (1D: (2 + cos t) / 3)
            o.pressure_mass, o.mass_its = self.mass.h, self.mass_its
            o.mass_lmin, o.mass_lmax = 3.0 ** -self.space.dim * 0.95, 1.0
        else:
            o.pressure_mass, o.mass_its, o.mass_lmin, o.mass_lmax = None, 0, 0.0, 1.0
        for i, m in enumerate(self.component_mgs):
            o.component_mg[i] = m.h.value
        return o


def solveStokesFGMRES(A: SparseMatrix, prec: StokesBlockPreconditioner, b: DenseVector, x: DenseVector,
                      rtol: float, restart: int = 50, maxit: int = 1000):
    """FGMRES(restart) on the Stokes system to ||b - A x|| <= rtol ||b|| (x: start iterate in,
    solution out); returns (info, relative residual estimate per iteration)"""
    info = N.SolveInfo()
    hist = np.zeros(max(maxit, 1))
    o = prec.opts(restart, maxit)
    check(N.lib().jfem_solve_stokes_fgmres(A.h, prec.mg.h, prec.scale.h, C.byref(o), b.h, x.h, rtol,
                                           C.byref(info), dptr(hist)))
    return info, hist[:info.iterations]
