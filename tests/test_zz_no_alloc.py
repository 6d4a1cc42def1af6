"""jfem_smooth_noalloc (the smoother entry point that allocates nothing) against jfem_smooth,
both called through the C ABI: Jacobi, Richardson and Chebyshev-Jacobi on two matrices in a
row, the second smaller than the first (the static scratch keeps the first call's contents),
bit for bit (the same kernels in the same order); a matrix above the stated maximum is refused
before any kernel runs."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _mat(jf, n, seed):
    """symmetric, diagonally dominant, about 15 entries per row (7 random columns per row,
    mirrored, plus the diagonal); generated row by row in O(n)"""
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), 7)
    cols = rng.integers(0, n, size=rows.shape[0])
    vals = 0.2 * rng.random(rows.shape[0])
    M = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    M = (M + M.T + sp.diags(2.0 + rng.random(n))).tocsr()
    M.sum_duplicates()
    M.sort_indices()
    return jf.SparseMatrix.from_csr(M.indptr, M.indices, M.data)


@pytest.mark.parametrize("kind,runs,omega,degree,lmin,lmax", [(1, 3, 0.6, 0, 0.0, 0.0), (0, 4, 0.05, 0, 0.0, 0.0),
                                                            (2, 2, 1.0, 4, 0.1, 2.0)])
def test_noalloc_smoother_matches_allocating_one(jf, kind, runs, omega, degree, lmin, lmax):
    from jfem_amd import _native as N
    L = N.lib()
    for call, n in enumerate((90000, 20011)):
        A = _mat(jf, n, 31 + call)
        rng = np.random.default_rng(7 + call)
        b = jf.DenseVector(rng.standard_normal(n))
        x0 = rng.standard_normal(n)
        xa, xb = jf.DenseVector(x0), jf.DenseVector(x0)
        assert L.jfem_smooth(A.h, kind, runs, omega, degree, lmin, lmax, b.h, xa.h) == 0, L.jfem_last_error()
        assert L.jfem_smooth_noalloc(A.h, kind, runs, omega, degree, lmin, lmax, b.h, xb.h) == 0, \
            L.jfem_last_error()
        ya, yb = xa.numpy(), xb.numpy()
        assert np.isfinite(ya).all()
        assert not np.array_equal(ya, x0)
        assert np.array_equal(ya, yb), (call, np.max(np.abs(ya - yb)))


def test_noalloc_smoother_refuses_oversize(jf):
    from jfem_amd import _native as N
    L = N.lib()
    n = 1048576 + 1
    idx = np.arange(n + 1, dtype=np.int64)
    A = jf.SparseMatrix.from_csr(idx, np.arange(n, dtype=np.int32), np.ones(n))
    b, x = jf.DenseVector(n), jf.DenseVector(n)
    assert L.jfem_smooth_noalloc(A.h, 1, 1, 1.0, 0, 0.0, 0.0, b.h, x.h) == 1   # JFEM_ERR_ARG
    assert b"JFEM_NOALLOC_MAX_ROWS" in L.jfem_last_error()
