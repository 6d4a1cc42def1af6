"""Smoother.smooth on any matrix (jfem_smooth through the C ABI) against a NumPy restatement
of the reference smoothers and against the closed forms that pin them:
  JacobiSmoother (Core/src/main/java/multigrid/JacobiSmoother.java:12-50): diagonal entries
      with |d| < 1e-10 are taken as 1, then `runs` times x = x + (D^-1 (b - A x)) omega;
  RichardsonSmoother (RichardsonSmoother.java:30-48): `runs` times r = b - A x, stop once
      ||r|| < 1e-10, x = x + r omega;
  Chebyshev-Jacobi (new; Saad, Alg. 12.1): each eigenvector v of D^-1 A with eigenvalue
      lambda is damped by T_m((theta - lambda) / delta) / T_m(theta / delta) per application."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _jacobi_np(A, b, x, runs, omega):
    d = A.diagonal().copy()
    d[np.abs(d) < 1e-10] = 1.0
    dinv = 1.0 / d
    x = x.copy()
    for _ in range(runs):
        r = b - A @ x
        x = x + (dinv * r) * omega
    return x


def _richardson_np(A, b, x, runs, omega):
    x = x.copy()
    sweeps = 0
    for _ in range(runs):
        r = b - A @ x
        if np.linalg.norm(r) < 1e-10:
            break
        x = x + r * omega
        sweeps += 1
    return x, sweeps


def _random_matrix(n, seed, tiny_diag_rows=()):
    """a nonsymmetric sparse matrix with a dominant diagonal and some diagonals below 1e-10"""
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=8.0 / n, random_state=seed, format="csr") * 0.3
    M = M + sp.diags(2.0 + rng.random(n))
    M = M.tolil()
    for i in tiny_diag_rows:
        M[i, i] = 3e-11
    return M.tocsr()


def _mat(jf, M):
    M = sp.csr_matrix(M)
    M.sort_indices()
    return jf.SparseMatrix.from_csr(M.indptr, M.indices, M.data)


@pytest.mark.parametrize("runs,omega", [(1, 1.0), (3, 0.6), (5, 0.8)])
def test_jacobi_smoother_matches_restatement(jf, runs, omega):
    """JacobiSmoother on a nonsymmetric matrix, including rows whose diagonal is below 1e-10
    (taken as 1): to 1e-13 relative"""
    n = 700
    M = _random_matrix(n, 3, tiny_diag_rows=(5, 333, 699))
    A = _mat(jf, M)
    rng = np.random.default_rng(4)
    b, x0 = rng.standard_normal(n), rng.standard_normal(n)
    x = jf.DenseVector(x0)
    jf.JacobiSmoother(runs, omega).smooth(A, jf.DenseVector(b), x)
    exp = _jacobi_np(M, b, x0, runs, omega)
    assert np.linalg.norm(x.numpy() - exp) <= 1e-13 * np.linalg.norm(exp)


def test_richardson_smoother_matches_restatement_and_exits(jf):
    """RichardsonSmoother: equal to the restatement on a generic system; with a residual of
    norm 1e-12 the iterate stays unchanged bit for bit; on 0.5 I with a residual of 1e-8 the
    sweep halves it each time and stops after 7 sweeps (||r|| < 1e-10), as the restatement does"""
    n = 500
    M = _random_matrix(n, 8)
    A = _mat(jf, M)
    rng = np.random.default_rng(9)
    b, x0 = rng.standard_normal(n), rng.standard_normal(n)
    x = jf.DenseVector(x0)
    jf.RichardsonSmoother(0.1, 4).smooth(A, jf.DenseVector(b), x)
    exp, sweeps = _richardson_np(M, b, x0, 4, 0.1)
    assert sweeps == 4
    assert np.linalg.norm(x.numpy() - exp) <= 1e-13 * np.linalg.norm(exp)

    d = rng.standard_normal(n)
    bb = M @ x0 + 1e-12 * d / np.linalg.norm(d)
    x = jf.DenseVector(x0)
    jf.RichardsonSmoother(0.1, 4).smooth(A, jf.DenseVector(bb), x)
    assert np.array_equal(x.numpy(), x0)

    half = sp.identity(n, format="csr")          # A = I, omega 0.5: r_k = 0.5^k r_0
    Ah = _mat(jf, half)
    x0 = rng.standard_normal(n)
    bh = x0 + 1e-8 * d / np.linalg.norm(d)
    x = jf.DenseVector(x0)
    jf.RichardsonSmoother(0.5, 20).smooth(Ah, jf.DenseVector(bh), x)
    exp, sweeps = _richardson_np(half, bh, x0, 20, 0.5)
    assert sweeps == 7
    assert np.linalg.norm(x.numpy() - exp) <= 1e-15 * np.linalg.norm(exp)


@pytest.mark.parametrize("deg,runs,lmin,lmax", [(4, 2, 0.1, 2.0), (3, 1, 0.3, 1.8)])
def test_chebyshev_smoother_closed_form(jf, oracle, deg, runs, lmin, lmax):
    """jfem_smooth's Chebyshev-Jacobi on the assembled 1D Q1 Laplacian (Dirichlet eliminated):
    each interior eigenvector sin(j pi i / n) of D^-1 A = tridiag(-1/2, 1, -1/2), eigenvalue
    1 - cos(j pi / n), comes out times the scaled Chebyshev polynomial, to 1e-12"""
    n = 32
    s = jf.ContinuousTPFESpace([0.0], [1.0], [n])
    s.assembleCells()
    s.assembleFunctions(1)
    s.initializeSystemMatrix()
    s.initializeRhs()
    s.evaluateCellIntegrals([jf.TPCellIntegral("GradGrad")], [])
    s.setBoundaryValues(None)
    A = s.getSystemMatrix()
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    c = [0] * deg + [1]
    i = np.arange(n + 1)
    sm = jf.ChebyshevSmoother(deg, runs, ratio=lmax / lmin)
    for j in (1, 5, 16, 27, 31):
        lam = 1 - np.cos(j * np.pi / n)
        f = (np.polynomial.chebyshev.chebval((theta - lam) / delta, c)
             / np.polynomial.chebyshev.chebval(theta / delta, c)) ** runs
        v = np.sin(j * np.pi * i / n)
        x = jf.DenseVector(v)
        sm.smooth(A, jf.DenseVector(np.zeros(n + 1)), x, lmax=lmax)
        assert np.linalg.norm(x.numpy() - f * v) <= 1e-12 * np.linalg.norm(v), j
