"""The OpenMP CPU baseline (oracle/cpu_baseline.c) computes what the oracle computes: its
prolongation matrices equal the oracle's (MGPreconditionerSpace.java:59-76), its Chebyshev
interval is the same power-iteration estimate, its V-cycle (mgStepV) and MG-preconditioned CG
(solvePCG) agree with the oracle's.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp


def _levels(O, base, k, nlev):
    G = O.MG(base, k, nlev, [0.0] * len(base), [1.0] * len(base), [O.GRAD_GRAD], [1.0], smoother=O.SM_CHEBYSHEV,
             runs=2, degree=4, lmin=[0.0] * nlev, lmax=[0.0] * nlev, cycles=1, fmg=False)
    mats = []
    for l in range(nlev):
        A = G.level_matrix(l)
        mats.append((A.rowptr, A.cols, A.vals))
    return G, mats


@pytest.mark.parametrize("base,k", [([3, 2], 1), ([2, 2], 2), ([2, 1, 2], 2), ([1, 2, 1], 1)])
def test_cpu_baseline_matches_oracle(oracle, base, k):
    from oracle import cpu_baseline as CB
    O = oracle
    nlev = 3
    G, mats = _levels(O, base, k, nlev)
    M = CB.CpuMG(len(base), k, base, mats)
    for l in range(nlev - 1):
        rp, ci, v = M.prolongation(l)
        nf = len(rp) - 1
        P = sp.csr_matrix((v, ci, rp), shape=(nf, mats[l][0].shape[0] - 1))
        Pr = O.prolongation([c << l for c in base], k, [0.0] * len(base), [1.0] * len(base)).to_scipy()
        assert P.nnz == Pr.nnz
        assert abs(P - Pr).max() <= 1e-15
    for l in range(1, nlev):
        lm = M.lmax(l)
        assert abs(lm - 1.1 * G.estimate_lmax(l)) <= 1e-12 * lm
        G.set_interval(l, lm / 20.0, lm)
    n = mats[-1][0].shape[0] - 1
    rng = np.random.default_rng(3)
    cells = [c << (nlev - 1) for c in base]
    b = rng.standard_normal(n)
    b[O.boundary_nodes(cells, k)] = 0
    x = 0.1 * rng.standard_normal(n)
    got, exp = M.vcycle(x, b), G.vcycle(x, b)
    assert np.linalg.norm(got - exp) <= 1e-12 * np.linalg.norm(exp)
    tol = 1e-9 * np.linalg.norm(b)
    xc, it, secs, res = M.pcg(b, tol)
    xo, ito, reso, _ = O.solve("cg", G.level_matrix(nlev - 1), b, np.zeros(n), tol, precond=G)
    A = G.level_matrix(nlev - 1).to_scipy()
    assert np.linalg.norm(b - A @ xc) <= 1.01 * tol
    assert abs(it - ito) <= 1
    assert np.linalg.norm(xc - xo) <= 1e-8 * np.linalg.norm(xo)


def test_cpu_baseline_reports_threads_and_model():
    from oracle import cpu_baseline as CB
    assert CB.threads() >= 1
    assert isinstance(CB.cpu_model(), str) and CB.cpu_model()
