"""The 3D Q2 heterogeneous-coefficient workload at reduced size: -div(a grad u) = f on 12^3
cells, a = exp(g) with g the seeded squared-exponential Gaussian random field (length 0.1)
rescaled to contrast 1e4, zero Dirichlet data, seeded 16-term sinusoid load.  The multigrid
hierarchy (3, 6, 12 cells per axis, Chebyshev-Jacobi V(2,2) of degree 4) is built with coarse
systems re-assembled from the children's mean coefficient (createSystem per level,
MGPreconditionerSpace.java:28-57) or as Galerkin products P^T A P (AMGFromMatrix,
MGPreconditionerInterface.java:387-471), in FP64 (inside CG) and in FP32 (inside FP64
flexible CG).  Every solve is checked on the oracle's assembled matrix: the true residual
meets the tolerance, and the distance to the LU solution obeys ||x - x_LU|| <= ||b - A x|| /
lambda_min(A), which holds for any x of an SPD system."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu

CELLS, K, NLEV = [12, 12, 12], 2, 3


@pytest.fixture(scope="module")
def c3_reference(oracle):
    from jfem_amd.workloads import gaussian_random_field, sinusoid_terms
    a = np.exp(gaussian_random_field(3, CELLS, 0.1, 1e4))
    ref = oracle.assemble(CELLS, K, [0, 0, 0], [1, 1, 1], [0], [1.0], cell_coeff=a)
    kw, amp = sinusoid_terms(3, 16, 3)
    rhs = oracle.rhs_sines(CELLS, K, [0, 0, 0], [1, 1, 1], kw, amp)
    oracle.dirichlet(ref, rhs, oracle.boundary_nodes(CELLS, K))
    A = ref.to_scipy().tocsc()
    lmin = spla.eigsh(A, k=1, sigma=0, which="LM", return_eigenvectors=False)[0]
    x_lu = spla.spsolve(A, rhs)
    return dict(a=a, kw=kw, amp=amp, A=A, rhs=rhs, lmin=lmin, x_lu=x_lu)


def _hierarchy(jf, a, precision, galerkin):
    spaces = []
    for l in range(NLEV):
        s = jf.ContinuousTPFESpace([0.0] * 3, [1.0] * 3, [c >> (NLEV - 1 - l) for c in CELLS])
        s.assembleCells()
        s.assembleFunctions(K)
        spaces.append(s)
    return jf.MGPreconditioner(spaces, [jf.TPCellIntegral("GradGrad")], jf.ChebyshevSmoother(4, 2), cycles=1,
                               fmg=False, precision=precision, cell_coeff=a, galerkin=galerkin), spaces


def test_c3_coarse_levels_reassembled(jf, oracle, c3_reference):
    """the re-assembled 6^3 level is the oracle's assembly with the mean of the 2^3 children's
    coefficients (pattern bit-exact, values to 1e-13 of the row scale)"""
    a = c3_reference["a"]
    G, _ = _hierarchy(jf, a, 0, False)
    ac = a.reshape(CELLS)
    mean = 0.125 * sum(ac[i::2, j::2, k::2] for i in (0, 1) for j in (0, 1) for k in (0, 1))
    cc = [c // 2 for c in CELLS]
    ref = oracle.assemble(cc, K, [0, 0, 0], [1, 1, 1], [0], [1.0], cell_coeff=mean.reshape(-1))
    oracle.dirichlet(ref, np.zeros(ref.n), oracle.boundary_nodes(cc, K))
    rp, cols, vals = G.getSystem(1).to_csr()
    assert np.array_equal(rp, ref.rowptr) and np.array_equal(cols, ref.cols)
    scale = np.repeat(np.maximum.reduceat(np.abs(ref.vals), ref.rowptr[:-1]), np.diff(ref.rowptr))
    assert np.max(np.abs(vals - ref.vals) / np.maximum(scale, 1e-300)) < 1e-13


@pytest.mark.parametrize("galerkin", [False, True])
def test_c3_fp32_multigrid_in_fp64_fcg(jf, c3_reference, galerkin):
    """FP64 MG-PCG and FP32 MG inside FP64 flexible CG, to ||r|| <= 1e-8 ||b||: the true FP64
    residual on the oracle's matrix meets the tolerance for both, both solutions lie within
    ||r|| / lambda_min of the LU solution, and the FP32-FP64 distance is reported"""
    R = c3_reference
    a, A, rhs, lmin, x_lu = R["a"], R["A"], R["rhs"], R["lmin"], R["x_lu"]
    tol = 1e-8 * np.linalg.norm(rhs)
    out = {}
    for precision in (0, 1):
        G, spaces = _hierarchy(jf, a, precision, galerkin)
        fine = spaces[-1]
        fine.initializeRhs()
        fine.writeCellIntegralsToRhs([jf.TPRightHandSideIntegral(jf.SinusoidSource(R["kw"], R["amp"]))],
                                     fine.getRhs())
        b = fine.getRhs().numpy()
        b[fine.boundary_nodes()] = 0
        # the device load sums shared DOFs in another order than the oracle: 1e-13 of max |rhs|
        assert np.max(np.abs(b - rhs)) <= 1e-13 * np.max(np.abs(rhs))
        it = jf.IterativeSolver()
        Af = G.getFinestSystem()
        bv = jf.DenseVector(rhs)
        if precision == 0:
            x = it.solvePCG(Af, G, bv, tol, x0=jf.DenseVector(np.zeros(len(rhs)))).numpy()
        else:
            x = it.solveFCG(Af, G, bv, tol).numpy()
        res = np.linalg.norm(rhs - A @ x)
        assert res <= 1.01 * tol, (precision, res / np.linalg.norm(rhs))
        err = np.linalg.norm(x - x_lu)
        assert err <= 1.001 * res / lmin, (precision, err, res / lmin)
        assert it.iterations <= 30, (precision, it.iterations)
        out[precision] = (x, it.iterations, res)
    d = np.linalg.norm(out[1][0] - out[0][0]) / np.linalg.norm(out[0][0])
    assert d <= (out[0][2] + out[1][2]) / lmin / np.linalg.norm(out[0][0]) * 1.001
    print(f"\nC3 12^3 {'Galerkin' if galerkin else 're-assembled'}: FP64 PCG {out[0][1]} it, rel.res "
          f"{out[0][2] / np.linalg.norm(rhs):.2e}; FP32 MG in FCG {out[1][1]} it, rel.res "
          f"{out[1][2] / np.linalg.norm(rhs):.2e}; |x32 - x64| / |x64| = {d:.2e}; "
          f"|x64 - x_LU| / |x_LU| = {np.linalg.norm(out[0][0] - x_lu) / np.linalg.norm(x_lu):.2e}")
