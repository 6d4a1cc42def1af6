"""MI355X (gfx950) implementation of JFEM's solve path for tensor-product Qk
spaces: structured space and DOF numbering, TPCellIntegral assembly into CSR,
Dirichlet elimination, SpMV, CG/BiCGStab/GMRES and geometric multigrid, all in
libjfem_hip.so behind the C ABI of include/jfem_hip.h."""
from ._native import (IllegalArgumentException, IllegalStateException, JfemError, NativeLibraryError,
                      UnsupportedOperationException, lib)
from .api import (FP32, FP64, GMRES2, MGSpacePreconditioner, ChebyshevSmoother, ContinuousTPFESpace, Context, DenseVector, IterativeSolver,
                  JacobiSmoother, MGPreconditioner, QuadratureRule1D, RichardsonSmoother, ScalarFESpaceFunction,
                  SinusoidSource, SmootherNet, SparseMatrix, TPCellIntegral, TPRightHandSideIntegral, context,
                  normL2Difference)

from .mixed import (MixedCellIntegral, MixedRightHandSideIntegral, MixedTPCellIntegral, StokesBlockPreconditioner,
                    TaylorHoodSpace, TPVectorCellIntegral, TPVectorRightHandSideIntegral, TrigSource, VectorTrigSource,
                    solveStokesFGMRES)

__all__ = [
    "IllegalArgumentException", "IllegalStateException", "JfemError", "NativeLibraryError",
    "UnsupportedOperationException", "lib", "FP32", "FP64", "GMRES2", "MGSpacePreconditioner", "ChebyshevSmoother", "ContinuousTPFESpace", "Context",
    "DenseVector", "IterativeSolver", "JacobiSmoother", "MGPreconditioner", "QuadratureRule1D",
    "RichardsonSmoother", "ScalarFESpaceFunction", "SinusoidSource", "SmootherNet", "SparseMatrix", "TPCellIntegral",
    "TPRightHandSideIntegral", "context", "normL2Difference", "MixedCellIntegral", "MixedRightHandSideIntegral",
    "MixedTPCellIntegral", "StokesBlockPreconditioner", "TaylorHoodSpace", "TPVectorCellIntegral",
    "TPVectorRightHandSideIntegral", "TrigSource", "VectorTrigSource", "solveStokesFGMRES",
]
