"""Pins the CPU oracle (oracle/jfem_oracle.c) against the reference's golden
vectors, closed forms and brute force.  CPU only."""
import itertools

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

# ---------------------------------------------------------------- golden arrays
# ContinuousIntegralTest.testTPCellIntegralDeg1G2 (Core/src/test/java/tensorproduct/
# ContinuousIntegralTest.java:23-40): Q1 cell [0,1]x[2.3,3.1], Gauss2, 4 decimals.
GG = [0.6833, -0.0583, -0.2833, -0.3417, -0.0583, 0.6833, -0.3417, -0.2833, -0.2833, -0.3417,
      0.6833, -0.0583, -0.3417, -0.2833, -0.0583, 0.6833]
GGW = [7.7812, -0.4688, -3.1612, -4.1513, -0.4688, 8.3437, -4.1513, -3.7237, -3.1612, -4.1513,
       8.2612, -0.9488, -4.1513, -3.7237, -0.9488, 8.8238]
VV = [0.0889, 0.0444, 0.0444, 0.0222, 0.0444, 0.0889, 0.0222, 0.0444,
      0.0444, 0.0222, 0.0889, 0.0444, 0.0222, 0.0444, 0.0444, 0.0889]
VVW = [0.9444, 0.5000, 0.5100, 0.2700, 0.5000, 1.0556, 0.2700, 0.5700, 0.5100, 0.2700, 1.0956,
       0.5800, 0.2700, 0.5700, 0.5800, 1.2244]
VG = [-0.3625, 0.1486, -0.0222, 0.2361, -0.3181, -0.1097, 0.0001, 0.4277, -0.2472, 0.0111, -0.1428,
      0.3789, -0.2249, -0.2473, -0.0984, 0.5706]
GV = [-0.3625, -0.3181, -0.2472, -0.2249, 0.1486, -0.1097, 0.0111, -0.2473, -0.0222, 0.0001,
      -0.1428, -0.0984, 0.2361, 0.4277, 0.3789, 0.5706]


def test_golden_cell_integrals(oracle):
    O = oracle
    s, e = [0, 2.3], [1, 3.1]

    def loc(i):
        return [i % 2, i // 2]

    def w(x):
        return x[1] * (x[0] + 4)

    def vw(x):
        return [2 - x[0] + x[1] * 0.1, x[0] * x[1]]

    k = 0
    for i in range(4):
        for j in range(4):
            assert abs(O.cell_integral(s, e, 1, O.GRAD_GRAD, 2, loc(i), loc(j)) - GG[k]) <= 6e-5
            assert abs(O.cell_integral(s, e, 1, O.GRAD_GRAD, 2, loc(i), loc(j), wfun=w) - GGW[k]) <= 6e-5
            assert abs(O.cell_integral(s, e, 1, O.VALUE_VALUE, 2, loc(i), loc(j)) - VV[k]) <= 6e-5
            assert abs(O.cell_integral(s, e, 1, O.VALUE_VALUE, 2, loc(i), loc(j), wfun=w) - VVW[k]) <= 6e-5
            assert abs(O.cell_integral(s, e, 1, O.VALUE_GRAD, 2, loc(i), loc(j), vfun=vw) - VG[k]) <= 6e-5
            assert abs(O.cell_integral(s, e, 1, O.GRAD_VALUE, 2, loc(i), loc(j), vfun=vw) - GV[k]) <= 6e-5
            k += 1


def test_golden_detects_transposed_operands(oracle):
    """GRAD_VALUE is not symmetric: swapping u and v must break the golden array."""
    O = oracle
    vw = lambda x: [2 - x[0] + x[1] * 0.1, x[0] * x[1]]
    dev = max(abs(O.cell_integral([0, 2.3], [1, 3.1], 1, O.GRAD_VALUE, 2, [j % 2, j // 2], [i % 2, i // 2],
                                  vfun=vw) - GV[i * 4 + j]) for i in range(4) for j in range(4))
    assert dev > 1e-2


# ---------------------------------------------------------------- quadrature / basis
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_gauss_exactness(oracle, n):
    p, w = oracle.quad(n)
    assert len(p) == n
    for deg in range(2 * n):
        assert abs(np.dot(w, p ** deg) - 1.0 / (deg + 1)) < 1e-14
    assert abs(np.dot(w, p ** (2 * n)) - 1.0 / (2 * n + 1)) > 1e-6


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6])
def test_gauss_lobatto(oracle, n):
    p, w = oracle.quad(10 + n)
    assert p[0] == 0.0 and p[-1] == 1.0
    # the reference lists GaussLobatto2's weights as {1/2, 1/2} on [-1,1]
    # (QuadratureRule1D.java:42) and halves them again in the constructor, so
    # that rule integrates to 1/2 of the true value; the oracle keeps the quirk.
    scale = 0.5 if n == 2 else 1.0
    for deg in range(2 * n - 2):
        assert abs(np.dot(w, p ** deg) - scale / (deg + 1)) < 1e-14


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_lagrange_basis(oracle, k):
    O = oracle
    for t in np.linspace(0, 1, 13):
        assert abs(sum(O.lagrange_value(t, l, k) for l in range(k + 1)) - 1) < 1e-13
        assert abs(sum(O.lagrange_deriv(t, l, k) for l in range(k + 1))) < 1e-11
    for l in range(k + 1):
        for m in range(k + 1):
            assert abs(O.lagrange_value(m / k, l, k) - (1.0 if l == m else 0.0)) < 1e-14
        # derivative against a central difference of the value
        for t in (0.1, 0.37, 0.8):
            h = 1e-6
            fd = (O.lagrange_value(t + h, l, k) - O.lagrange_value(t - h, l, k)) / (2 * h)
            assert abs(fd - O.lagrange_deriv(t, l, k)) < 1e-6


# ---------------------------------------------------------------- element matrices
def k1d(k, h):
    if k == 1:
        return np.array([[1, -1], [-1, 1]]) / h, np.array([[2, 1], [1, 2]]) * h / 6
    return (np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * h),
            np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) * h / 30)


def tensor_element(k, hs, kinds_weights):
    """sum over integrals of the Kronecker form; local index with axis 0 fastest."""
    d = len(hs)
    KM = [k1d(k, h) for h in hs]
    out = 0
    for kind, w in kinds_weights:
        if kind == 0:      # GRAD_GRAD = sum_a (K on axis a, M elsewhere)
            for a in range(d):
                mats = [KM[b][0] if b == a else KM[b][1] for b in range(d)]
                m = mats[d - 1]
                for b in range(d - 2, -1, -1):
                    m = np.kron(m, mats[b])
                out = out + w * m
        else:
            mats = [KM[b][1] for b in range(d)]
            m = mats[d - 1]
            for b in range(d - 2, -1, -1):
                m = np.kron(m, mats[b])
            out = out + w * m
    return out


@pytest.mark.parametrize("d,k", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_element_matrix_closed_form(oracle, d, k):
    cells = [3, 2, 2][:d]
    start = [0.0, -1.0, 0.5][:d]
    end = [1.5, 1.0, 2.0][:d]
    hs = [(e - s) / c for s, e, c in zip(start, end, cells)]
    E = oracle.element_matrix(cells, k, start, end, 1, [0, 1], [1.0, 5.0])
    ref = tensor_element(k, hs, [(0, 1.0), (1, 5.0)])
    assert np.max(np.abs(E - ref)) < 1e-13 * np.max(np.abs(ref))


# ---------------------------------------------------------------- DOF numbering
def closed_form_dof(cells, k, p):
    """prefix(owner cell) + rank, written from the first-appearance rule."""
    d = len(cells)
    T = [k * n + 1 for n in cells]
    pre, mprod, rank = 0, 1, 0
    for a in range(d):
        c = 0 if p[a] == 0 else (p[a] - 1) // k
        l = p[a] - c * k
        tail = int(np.prod(T[a + 1:])) if a + 1 < d else 1
        S = c * k + 1 if c > 0 else 0
        pre += mprod * S * tail
        rank += (l - (1 if c > 0 else 0)) * mprod
        mprod *= k if c > 0 else k + 1
    return pre + rank


@pytest.mark.parametrize("cells,k", [([1, 1], 1), ([2, 3], 1), ([3, 2], 2), ([3, 2, 2], 2), ([2, 2, 3], 1),
                                     ([4], 3), ([2, 3], 3), ([2, 1, 2], 4)])
def test_dof_numbering_closed_form(oracle, cells, k):
    n, table = oracle.dof_table(cells, k)
    assert n == int(np.prod([k * c + 1 for c in cells]))
    d = len(cells)
    for cell in range(table.shape[0]):
        cc = np.unravel_index(cell, cells)          # last axis fastest
        for li in range(table.shape[1]):
            l = [(li // (k + 1) ** a) % (k + 1) for a in range(d)]
            p = [cc[a] * k + l[a] for a in range(d)]
            assert table[cell, li] == closed_form_dof(cells, k, p)


def test_dof_numbering_first_cell_is_lexicographic(oracle):
    n, t = oracle.dof_table([3, 2], 2)
    assert list(t[0]) == list(range(9))
    assert list(t[1][:3]) == [6, 7, 8]      # shared face of cells (0,0) and (0,1)
    assert t[2][0] == 2                      # cell (1,0) starts on the shared face


# ---------------------------------------------------------------- assembly
def structural_nnz(cells, k):
    per_axis = []
    for n in cells:
        cnt = 0
        for p in range(k * n + 1):
            if p % k == 0:
                lo, hi = max(p // k - 1, 0), min(p // k, n - 1)
            else:
                lo = hi = p // k
            cnt += k * (hi - lo + 1) + 1
        per_axis.append(cnt)
    return int(np.prod(per_axis))


@pytest.mark.parametrize("cells,k", [([8, 8], 1), ([5, 4], 2), ([3, 3, 3], 1), ([3, 2, 3], 2)])
def test_assembly_invariants(oracle, cells, k):
    O = oracle
    d = len(cells)
    start, end = [0.0] * d, [1.0, 2.0, 0.5][:d]
    A = O.assemble(cells, k, start, end, [O.GRAD_GRAD], [1.0])
    M = O.assemble(cells, k, start, end, [O.VALUE_VALUE], [1.0])
    S = A.to_scipy()
    assert A.nnz == structural_nnz(cells, k)
    assert abs(S - S.T).max() == 0.0                              # bit-symmetric
    assert np.max(np.abs(S.sum(axis=1))) < 1e-12 * abs(S).max()   # constants are in the kernel
    vol = np.prod(np.array(end) - np.array(start))
    assert abs(M.to_scipy().sum() - vol) < 1e-12 * vol            # partition of unity integrates to |Omega|
    # columns sorted within rows
    for r in range(A.n):
        c = A.cols[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(c) > 0)


def test_assembly_2d_q2_keeps_rounding_level_entries(oracle):
    """2D Q2 GRAD_GRAD has element entries that vanish mathematically
    (K02*M11 + M02*K11 = 16/90 - 16/90).  Evaluated by the reference's
    quadrature they come out at rounding level, not exactly 0.0, so
    addToMatrix keeps them and the pattern stays structural."""
    O = oracle
    E = O.element_matrix([4, 4], 2, [0, 0], [1, 1], 0, [O.GRAD_GRAD], [1.0])
    tiny = np.abs(E)[np.abs(E) < 1e-10]
    assert len(tiny) == 4 and np.all(tiny > 0)
    A = O.assemble([4, 4], 2, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0])
    assert A.nnz == structural_nnz([4, 4], 2)


def test_assembly_matches_global_quadrature(oracle):
    """u^T A u equals the exact energy of a smooth function's interpolant."""
    O = oracle
    cells, k = [6, 5], 2
    A = O.assemble(cells, k, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0])
    xyz = O.node_coords(cells, k, [0, 0], [1, 1])
    u = xyz[:, 0] ** 2 + xyz[:, 0] * xyz[:, 1]     # in Q2, so interpolation is exact
    energy = u @ A.matvec(u)
    # int_0^1 int_0^1 |grad u|^2 = int (2x+y)^2 + x^2 = 4/3 + 1 + 1/3 + 1/3 = 3
    assert abs(energy - 3.0) < 1e-12


# ---------------------------------------------------------------- Dirichlet
def test_dirichlet_matches_dense_elimination(oracle):
    O = oracle
    cells, k = [4, 3], 2
    A = O.assemble(cells, k, [0, 0], [1, 1], [O.GRAD_GRAD, O.VALUE_VALUE], [1.0, 5.0])
    dense = A.to_scipy().toarray()
    rng = np.random.default_rng(1)
    rhs = rng.standard_normal(A.n)
    nodes = O.boundary_nodes(cells, k)
    g = rng.standard_normal(len(nodes))
    # symmetric elimination written densely
    free = np.setdiff1d(np.arange(A.n), nodes)
    exp_rhs = rhs.copy()
    exp_rhs[free] -= dense[np.ix_(free, nodes)] @ g
    exp_rhs[nodes] = g
    exp = dense.copy()
    exp[nodes, :] = 0
    exp[:, nodes] = 0
    exp[nodes, nodes] = 1
    r2 = rhs.copy()
    O.dirichlet(A, r2, nodes, g)
    assert np.allclose(A.to_scipy().toarray(), exp, rtol=0, atol=0)
    assert np.max(np.abs(r2 - exp_rhs)) < 1e-12 * np.max(np.abs(exp_rhs))
    assert A.nnz == int(np.count_nonzero(dense) + 0)   # explicit zeros stay in the pattern


def test_boundary_nodes(oracle):
    O = oracle
    cells, k = [3, 4], 2
    xyz = O.node_coords(cells, k, [0, 0], [1, 1])
    nodes = set(O.boundary_nodes(cells, k).tolist())
    on = {i for i in range(len(xyz)) if np.any(np.isclose(xyz[i], 0)) or np.any(np.isclose(xyz[i], 1))}
    assert nodes == on


# ---------------------------------------------------------------- solvers
def csr_of(dense):
    S = sp.csr_matrix(np.asarray(dense, float))
    from oracle.oracle import CSR
    return CSR(S.shape[0], S.indptr.astype(np.int64), S.indices.astype(np.int32), S.data.astype(float))


SYMM = [[0, 1, 2], [1, 0, 3], [2, 3, 0]]
NONSYMM = [[3, 1, 7], [2, 0, 3], [2, 3, 0]]
GOOD = [[3, 1, 7], [2, 4, 3], [2, 3, 7]]


@pytest.mark.parametrize("mat", [SYMM, NONSYMM, GOOD])
def test_bicgstab_known_answers(oracle, mat):
    """BiCGStabTest.java:19-40: A sol = b and sol = A^-1 b."""
    A = csr_of(mat)
    b = np.array([3.0, 4.0, 5.0])
    x0 = np.array([1.0, 0, 0])
    x, it, res, fb = oracle.solve("bicgstab", A, b, x0, 1e-12)
    assert np.allclose(np.array(mat) @ x, b, atol=1e-9)
    assert np.allclose(x, np.linalg.solve(mat, b), atol=1e-9)


@pytest.mark.parametrize("mat", [SYMM, NONSYMM, GOOD])
def test_gmres_known_answers(oracle, mat):
    A = csr_of(mat)
    b = np.array([3.0, 4.0, 5.0])
    x, it, res, fb = oracle.solve("gmres", A, b, np.array([1.0, 0, 0]), 1e-12)
    assert np.allclose(x, np.linalg.solve(mat, b), atol=1e-9)


@pytest.mark.parametrize("mat", [SYMM, NONSYMM, GOOD])
@pytest.mark.parametrize("precond", [None, "identity"])
def test_gmres2_known_answers(oracle, mat, precond):
    """GMResTest.testSymm / testNonSymm / testPSymm / testPNonSymm / testGoodPNonSymm
    (Core/src/test/java/linalg/GMResTest.java:20-40, 69-125): GMRES2(1e-12), start e_0"""
    A = csr_of(mat)
    b = np.array([3.0, 4.0, 5.0])
    x, it, res, fb = oracle.solve("gmres2", A, b, np.array([1.0, 0, 0]), 1e-12, precond=precond)
    assert np.allclose(np.array(mat) @ x, b, atol=1e-10)
    assert np.allclose(x, np.linalg.solve(mat, b), atol=1e-9)


def test_gmres2_identity_and_large(oracle):
    """GMResTest.testIdentity / testPIdentity (:11-18, 60-67) and testLarge / testPLarge
    (:42-58, 94-110) with a seeded matrix of the same construction"""
    n = 15
    b = np.arange(1.0, n + 1)
    e0 = np.zeros(n)
    e0[0] = 1
    for pc in (None, "identity"):
        x, it, res, fb = oracle.solve("gmres2", csr_of(np.eye(n)), b, e0, 1e-12, precond=pc)
        assert np.allclose(x, b, atol=1e-12)
    import scipy.sparse as sp
    n = 2000
    rng = np.random.default_rng(4)
    rows = np.concatenate([np.arange(10 * n) % n, rng.integers(0, n, 10 * n)])
    cols = np.concatenate([np.arange(10 * n) % n, rng.integers(0, n, 10 * n)])
    vals = np.concatenate([np.full(10 * n, 0.2), rng.random(10 * n)])
    M = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    M.sum_duplicates()
    M.sort_indices()
    A = oracle.CSR(n, M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data.copy())
    b = np.full(n, 10.0)
    e0 = np.zeros(n)
    e0[0] = 1
    x, it, res, fb = oracle.solve("gmres2", A, b, e0, 1e-10)
    assert np.linalg.norm(M @ x - b) <= 1.01e-10


@pytest.mark.parametrize("n", [5, 15])
def test_identity_solves(oracle, n):
    A = csr_of(np.eye(n))
    b = np.arange(1.0, n + 1)
    for m in ("cg", "bicgstab", "gmres"):
        x, *_ = oracle.solve(m, A, b, np.eye(n)[0], 1e-12)
        assert np.allclose(x, b, atol=1e-12)


def test_cg_on_fe_system_matches_lu(oracle):
    O = oracle
    cells, k = [12, 10], 1
    A = O.assemble(cells, k, [0, 0], [1, 1], [O.GRAD_GRAD, O.VALUE_VALUE], [1.0, 5.0])
    rhs = O.rhs_sines(cells, k, [0, 0], [1, 1], [[1, 2], [3, 1]], [1.0, -0.5])
    O.dirichlet(A, rhs, O.boundary_nodes(cells, k))
    x0 = np.zeros(A.n)
    x0[0] = 1
    x, it, res, _ = O.solve("cg", A, rhs, x0, 1e-12 * np.linalg.norm(rhs))
    exact = sla.lu_solve(sla.lu_factor(A.to_scipy().toarray()), rhs)
    assert np.linalg.norm(x - exact) <= 1e-10 * np.linalg.norm(exact)
    assert res <= 1e-12 * np.linalg.norm(rhs) * 1.0001


def test_lu_solve(oracle):
    rng = np.random.default_rng(3)
    M = rng.standard_normal((40, 40))
    b = rng.standard_normal(40)
    assert np.allclose(oracle.lu_solve(M, b), np.linalg.solve(M, b), atol=1e-10)


# ---------------------------------------------------------------- multigrid
def test_prolongation_reproduces_linear_functions(oracle):
    """MultiGridSpaceTest.testProlongateInterpolate: P v of a linear v equals v."""
    O = oracle
    for k, cc in ((1, [4, 4]), (2, [3, 2]), (2, [2, 2, 2])):
        d = len(cc)
        start, end = [100.7, 0.2, 0.0][:d], [1000.0, 0.21, 1.0][:d]
        P = O.prolongation(cc, k, start, end)
        xc = O.node_coords(cc, k, start, end)
        xf = O.node_coords([2 * c for c in cc], k, start, end)
        coef = np.array([0.3, 2.0, -1.0][:d])
        vc = xc @ coef + 1.5
        vf = xf @ coef + 1.5
        assert np.max(np.abs(P.matvec(vc) - vf)) <= 1e-8 * np.max(np.abs(vf))
        quad = lambda x: x[:, 0] ** 2 if k == 2 else x[:, 0]
        assert np.max(np.abs(P.matvec(quad(xc)) - quad(xf))) <= 1e-8 * np.max(np.abs(quad(xf)))


def test_galerkin_identity(oracle):
    """MultiGridSpaceTest: P^T A_fine P equals A_coarse for nested spaces."""
    O = oracle
    start, end = [100.7, 0.2], [1000.0, 0.21]
    for k in (1, 2):
        Ac = O.assemble([4, 4], k, start, end, [O.GRAD_GRAD], [1.0]).to_scipy()
        Af = O.assemble([8, 8], k, start, end, [O.GRAD_GRAD], [1.0]).to_scipy()
        P = O.prolongation([4, 4], k, start, end).to_scipy()
        G = (P.T @ Af @ P).toarray()
        assert np.max(np.abs(G - Ac.toarray())) <= 1e-9 * np.max(np.abs(Ac.toarray()))


def test_mg_preconditioned_gmres_laplace(oracle):
    """The LaplaceMG.testMG bound (Core/src/test/java/examples/LaplaceMG.java:23-101: Q1 on
    [0,1]^2, Richardson(0.1, 8) smoothing, MG-preconditioned GMRES to 1e-8 in fewer than 10
    iterations) with the MGPreconditionerSpace application (copy boundary values, full-V
    correction, 3 V-cycles); the test as written (MGSpace) is
    test_mgspace_preconditioned_gmres_laplace."""
    O = oracle
    for refinements in (2, 3):
        nlev = refinements + 1
        G = O.MG([4, 4], 1, nlev, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0], smoother=O.SM_RICHARDSON, runs=8,
                 omega=0.1, cycles=3, fmg=True)
        cells = [4 << refinements] * 2
        A = G.level_matrix(nlev - 1)
        rhs = O.rhs_sines(cells, 1, [0, 0], [1, 1], [[1, 1]], [1.0])
        rhs[O.boundary_nodes(cells, 1)] = 0
        x, it, res, _ = O.solve("gmres", A, rhs, np.zeros(A.n), 1e-8, precond=G)
        assert it < 10
        assert res <= 1e-8


def test_mgspace_preconditioned_gmres_laplace(oracle):
    """LaplaceMG.testMG as written (Core/src/test/java/examples/LaplaceMG.java:23-101): MGSpace
    with Richardson(0.1, 8) smoothers, preconditioner MGPReconditioner2 (v + fullVCycleSolver(v - A v),
    raw prolongation matrices), solvePGMRES to 1e-8 in fewer than 10 iterations; the application
    is linear (it must be, for GMRES) and differs from the MGPreconditionerSpace one"""
    O = oracle
    for refinements in (2, 3):
        nlev = refinements + 1
        G = O.MG([4, 4], 1, nlev, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0], smoother=O.SM_RICHARDSON, runs=8,
                 omega=0.1, cycles=1, fmg=2)
        G.set_raw(True)
        cells = [4 << refinements] * 2
        A = G.level_matrix(nlev - 1)
        rhs = O.rhs_sines(cells, 1, [0, 0], [1, 1], [[1, 1]], [1.0])
        rhs[O.boundary_nodes(cells, 1)] = 0
        x, it, res, _ = O.solve("gmres", A, rhs, np.zeros(A.n), 1e-8, precond=G)
        assert it < 10
        assert res <= 1e-8
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(A.n), rng.standard_normal(A.n)
    assert np.allclose(G.apply(2 * u - 3 * v), 2 * G.apply(u) - 3 * G.apply(v), atol=1e-9)
    H = O.MG([4, 4], 1, nlev, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0], smoother=O.SM_RICHARDSON, runs=8,
             omega=0.1, cycles=1, fmg=True)
    assert np.linalg.norm(G.apply(u) - H.apply(u)) > 1e-6 * np.linalg.norm(H.apply(u))


@pytest.mark.parametrize("sm", ["jacobi", "chebyshev"])
def test_mg_vcycle_contracts(oracle, sm):
    O = oracle
    kind = O.SM_JACOBI if sm == "jacobi" else O.SM_CHEBYSHEV
    nlev = 4
    G = O.MG([2, 2], 2, nlev, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0], smoother=kind, runs=2, omega=0.6,
             degree=4, lmin=[0.05] * nlev, lmax=[2.0] * nlev, cycles=1, fmg=False)
    A = G.level_matrix(nlev - 1).to_scipy()
    rng = np.random.default_rng(0)
    b = rng.standard_normal(A.shape[0])
    b[O.boundary_nodes([16, 16], 2)] = 0
    x = np.zeros_like(b)
    r0 = np.linalg.norm(b)
    for _ in range(5):
        x = G.vcycle(x, b)
    assert np.linalg.norm(b - A @ x) < 1e-3 * r0


@pytest.mark.parametrize("cells,k", [([3, 2], 1), ([2, 3], 2), ([2, 1, 2], 1)])
def test_assembly_grad_value_invariants(oracle, cells, k):
    """GRAD_VALUE (grad u . w) v and VALUE_GRAD (grad v . w) u with a constant vector w: for
    the interpolant of a linear f(x) = a . x (exact in Qk), A_GV f = (a . w) M 1 (the
    VALUE_VALUE matrix times ones), A_GV 1 = 0, and VALUE_GRAD is the transpose of GRAD_VALUE"""
    O = oracle
    d = len(cells)
    start, end = [0.0, -1.0, 0.5][:d], [1.0, 0.5, 2.0][:d]
    w = [0.3, -1.2, 0.7][:d]
    a = np.array([1.5, -0.4, 2.0][:d])
    GV = O.assemble(cells, k, start, end, [O.GRAD_VALUE], [w]).to_scipy()
    VG = O.assemble(cells, k, start, end, [O.VALUE_GRAD], [w]).to_scipy()
    M = O.assemble(cells, k, start, end, [O.VALUE_VALUE], [1.0]).to_scipy()
    x = O.node_coords(cells, k, start, end)
    f = x @ a
    one = np.ones(GV.shape[0])
    assert np.allclose(GV @ f, (a @ w) * (M @ one), rtol=0, atol=1e-12)
    assert np.allclose(GV @ one, 0, atol=1e-13)
    assert abs(GV - VG.T).max() <= 1e-15


def test_mg_wcycle_contracts_more_than_vcycle(oracle):
    """mgStepW's second coarse correction: one W-cycle leaves a smaller residual than one
    V-cycle from the same start (only slightly: the smoothing, shared by both, leaves the
    high frequencies that dominate the residual); on two levels, where the coarse V-cycle
    is the exact coarse solve, the W-cycle equals pre-smoothing, two exact two-grid
    corrections (restriction P^T, prolongation P, zero boundary values), post-smoothing,
    composed here from scipy and the oracle's prolongation matrix"""
    O = oracle
    nlev = 4
    G = O.MG([2, 2], 2, nlev, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0], smoother=O.SM_JACOBI, runs=2, omega=0.6,
             cycles=1, fmg=False)
    A = G.level_matrix(nlev - 1).to_scipy()
    rng = np.random.default_rng(1)
    b = rng.standard_normal(A.shape[0])
    b[O.boundary_nodes([16, 16], 2)] = 0
    x0 = np.zeros_like(b)
    rv = np.linalg.norm(b - A @ G.vcycle(x0, b))
    rw = np.linalg.norm(b - A @ G.wcycle(x0, b))
    assert rw < rv, (rw, rv)
    # on two levels the coarse V-cycle is the exact coarse solve, so the W-cycle equals
    # pre-smoothing, two exact two-grid corrections, post-smoothing
    G2 = O.MG([4, 4], 2, 2, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0], smoother=O.SM_JACOBI, runs=2, omega=0.6,
              cycles=1, fmg=False)
    A2 = G2.level_matrix(1).to_scipy()
    Ac = G2.level_matrix(0).to_scipy().toarray()
    P = O.prolongation([4, 4], 2, [0, 0], [1, 1]).to_scipy()
    bnd_f = O.boundary_nodes([8, 8], 2)
    bnd_c = O.boundary_nodes([4, 4], 2)
    b2 = rng.standard_normal(A2.shape[0])
    b2[bnd_f] = 0
    x = G2.smooth(1, b2, np.zeros_like(b2))
    for _ in range(2):
        rc = P.T @ (b2 - A2 @ x)
        rc[bnd_c] = 0
        e = P @ np.linalg.solve(Ac, rc)
        e[bnd_f] = 0
        x = x + e
    x = G2.smooth(1, b2, x)
    got = G2.wcycle(np.zeros_like(b2), b2)
    assert np.linalg.norm(got - x) <= 1e-12 * np.linalg.norm(x)


@pytest.mark.parametrize("cells,k", [([3, 2], 1), ([3, 2], 2), ([2, 2, 2], 1), ([2, 1, 2], 2)])
def test_fe_value_reproduces_qk_polynomials(oracle, cells, k):
    """Qk interpolation is exact on Qk: with the nodal values of a polynomial of
    degree <= k per axis as coefficients, the FE function equals the polynomial
    everywhere in the domain (on faces and nodes too), and is 0 outside it."""
    O = oracle
    d = len(cells)
    start, end = [0.0, -1.0, 0.25][:d], [1.0, 1.0, 2.0][:d]
    xyz = O.node_coords(cells, k, start, end)

    def u(p):
        v = 1.0 + 0.5 * p[:, 0] - 2.0 * p[:, 1]
        mono = np.ones(len(p))
        for a in range(d):
            mono = mono * p[:, a] ** k
        return v + 0.75 * mono

    rng = np.random.default_rng(3)
    pts = np.column_stack([rng.uniform(start[a], end[a], 25) for a in range(d)])
    # a node and a face point
    pts = np.vstack([pts, xyz[len(xyz) // 2], 0.5 * (xyz[1] + xyz[2])])
    got = O.fe_value(cells, k, start, end, u(xyz), pts)
    assert np.abs(got - u(pts)).max() <= 1e-12 * np.abs(u(pts)).max()
    outside = np.array([[end[a] + 0.5 for a in range(d)]])
    assert O.fe_value(cells, k, start, end, u(xyz), outside)[0] == 0.0


# ------------------------------------------- smoothers, pinned by their closed forms
# The 1D Q1 Laplacian on n uniform cells, after Dirichlet elimination
# (AcceptsMatrixBoundaryValues.java:18-44), is A = (1/h) tridiag(-1, 2, -1) on the
# interior nodes and the identity on the two boundary nodes.  D^-1 A restricted to the
# interior is tridiag(-1/2, 1, -1/2) with eigenpairs lambda_j = 1 - cos(j pi / n),
# v_j(i) = sin(j pi i / n), j = 1..n-1 (zero on the boundary nodes), and A's own
# eigenvalues are mu_j = 2 n lambda_j.  With b = 0 every smoother maps v_j to a known
# multiple of v_j:
#   Jacobi (JacobiSmoother.java:28-50), `runs` sweeps:          (1 - omega lambda)^runs
#   Richardson (RichardsonSmoother.java:30-48), `runs` sweeps:  (1 - omega mu)^runs
#   Chebyshev-Jacobi on [lmin, lmax] (Saad, Iterative Methods for Sparse Linear Systems,
#   2nd ed., Alg. 12.1), degree m, `runs` applications:
#       (T_m((theta - lambda) / delta) / T_m(theta / delta))^runs,
#       theta = (lmax + lmin) / 2, delta = (lmax - lmin) / 2.
def _laplace_1d(O, n, smoother, **kw):
    return O.MG([n], 1, 1, [0.0], [1.0], [O.GRAD_GRAD], [1.0], smoother=smoother, cycles=1, fmg=False, **kw)


def _cheb_factor(lam, lmin, lmax, deg):
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    c = [0] * deg + [1]
    return (np.polynomial.chebyshev.chebval((theta - lam) / delta, c)
            / np.polynomial.chebyshev.chebval(theta / delta, c))


def _smoother_deviation(O, G, n, factor):
    """max over the eigenvectors v_j of |smooth(0, v_j) - factor(lambda_j) v_j| / |v_j|"""
    i = np.arange(n + 1)
    dev = 0.0
    for j in range(1, n):
        lam = 1 - np.cos(j * np.pi / n)
        v = np.sin(j * np.pi * i / n)
        out = G.smooth(0, np.zeros(n + 1), v)
        dev = max(dev, np.linalg.norm(out - factor(lam) * v) / np.linalg.norm(v))
    return dev


def test_laplace_1d_spectrum(oracle):
    """the premise of the closed-form smoother tests: the oracle's level matrix and Jacobi
    diagonal are the ones stated above"""
    O = oracle
    n = 16
    G = _laplace_1d(O, n, O.SM_JACOBI, runs=1, omega=1.0)
    A = G.level_matrix(0).to_scipy().toarray()
    T = n * (2 * np.eye(n - 1) - np.eye(n - 1, k=1) - np.eye(n - 1, k=-1))
    assert np.allclose(A[1:-1, 1:-1], T, rtol=0, atol=1e-12 * n)
    assert A[0, 0] == 1 and A[n, n] == 1 and not A[0, 1:].any() and not A[1:, 0].any()
    dinv = G.level_dinv(0)
    assert np.allclose(dinv[1:-1], 1 / (2 * n), rtol=1e-14) and dinv[0] == 1 and dinv[n] == 1


@pytest.mark.parametrize("deg,runs,lmin,lmax", [(4, 1, 0.05, 2.0), (4, 2, 2.0 / 30, 2.0), (3, 2, 0.3, 1.8),
                                                (1, 3, 0.5, 2.0), (6, 1, 0.1, 2.2)])
def test_chebyshev_smoother_closed_form(oracle, deg, runs, lmin, lmax):
    """the oracle's Chebyshev-Jacobi smoother damps each eigenvector of D^-1 A by the scaled
    Chebyshev polynomial, to 1e-12; each planted mistake in the recurrence or the interval
    (c1 without rho_old, c2 without the factor 2, the centre taken as the half-width) breaks it"""
    O = oracle
    n = 24
    G = _laplace_1d(O, n, O.SM_CHEBYSHEV, runs=runs, degree=deg, lmin=[lmin], lmax=[lmax])
    f = lambda lam: _cheb_factor(lam, lmin, lmax, deg) ** runs
    assert _smoother_deviation(O, G, n, f) < 1e-12
    plants = [O.planted_smoother.CHEB_CENTRE] if deg == 1 else \
        [O.planted_smoother.CHEB_C1, O.planted_smoother.CHEB_C2, O.planted_smoother.CHEB_CENTRE]
    for plant in plants:     # degree 1 never reaches c1, c2
        with O.planted_smoother(plant):
            assert _smoother_deviation(O, G, n, f) > 1e-4, plant


@pytest.mark.parametrize("omega,runs", [(0.6, 2), (2.0 / 3.0, 3), (1.0, 1)])
def test_jacobi_smoother_closed_form(oracle, omega, runs):
    """(1 - omega lambda)^runs per eigenvector to 1e-12; dropping omega breaks it (omega != 1)"""
    O = oracle
    n = 24
    G = _laplace_1d(O, n, O.SM_JACOBI, runs=runs, omega=omega)
    f = lambda lam: (1 - omega * lam) ** runs
    assert _smoother_deviation(O, G, n, f) < 1e-12
    if omega != 1.0:
        with O.planted_smoother(O.planted_smoother.JACOBI_OMEGA):
            assert _smoother_deviation(O, G, n, f) > 1e-4


def test_richardson_smoother_closed_form_and_exit(oracle):
    """(1 - omega mu)^runs per eigenvector of A; the sweep stops once ||b - A x|| < 1e-10
    (RichardsonSmoother.java:38-40): with a residual of norm 1e-12 x stays unchanged bit for bit,
    and without that exit (planted) it moves"""
    O = oracle
    n, omega, runs = 12, 0.01, 3
    G = _laplace_1d(O, n, O.SM_RICHARDSON, runs=runs, omega=omega)
    f = lambda lam: (1 - omega * 2 * n * lam) ** runs
    assert _smoother_deviation(O, G, n, f) < 1e-12
    A = G.level_matrix(0).to_scipy()
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n + 1)
    d = rng.standard_normal(n + 1)
    b = A @ x + 1e-12 * d / np.linalg.norm(d)
    assert np.linalg.norm(b - A @ x) < 1e-10
    assert np.array_equal(G.smooth(0, b, x), x)
    with O.planted_smoother(O.planted_smoother.RICHARDSON_EXIT):
        assert not np.array_equal(G.smooth(0, b, x), x)


def test_face_quadrature_refused(oracle):
    """every integral kind but GRAD_GRAD evaluates the shape functions through value(x) /
    gradient(x), i.e. in the first support cell containing x (ScalarShapeFunction.java:26-39,
    TPCellIntegral.java:124-157, TPRightHandSideIntegral.java:26-29); Gauss-Lobatto points lie on
    cell faces, where that cell can be the neighbour, so the restatement (which integrates in the
    cell itself) refuses those kinds with Lobatto rules and keeps GRAD_GRAD (gradientInCell)"""
    O = oracle
    cells, k = [3, 2], 2
    for rule in (12, 13, 14, 15, 16):
        A = O.assemble(cells, k, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0], rule=rule)
        assert A.n == 35
        for kind in (O.VALUE_VALUE, O.H1):
            with pytest.raises(ValueError):
                O.assemble(cells, k, [0, 0], [1, 1], [kind], [1.0], rule=rule)
        with pytest.raises(ValueError):
            O.assemble(cells, k, [0, 0], [1, 1], [O.GRAD_GRAD, O.VALUE_VALUE], [1.0, 1.0], rule=rule)
        with pytest.raises(ValueError):
            O.rhs_sines(cells, k, [0, 0], [1, 1], [[1, 2]], [1.0], rule=rule)
    for rule in (1, 2, 3, 4, 5):
        O.assemble(cells, k, [0, 0], [1, 1], [O.VALUE_VALUE, O.H1], [1.0, 2.0], rule=rule)


def test_power_iteration_estimate_converges_to_lambda_max(oracle):
    """the oracle's power iteration on D^-1 A (the Chebyshev interval's upper end before the
    safety factor) converges to the largest eigenvalue of D^-1 A: within 1e-6 after 400
    iterations, and from below-or-near after 30 (the hierarchy's default)"""
    O = oracle
    G = O.MG([2, 2], 2, 3, [0, 0], [1, 1], [O.GRAD_GRAD], [1.0], smoother=O.SM_CHEBYSHEV, runs=1, degree=2,
             lmin=[0.0] * 3, lmax=[0.0] * 3, cycles=1, fmg=False)
    for l in (1, 2):
        A = G.level_matrix(l).to_scipy().toarray()
        lam = np.max(np.abs(np.linalg.eigvals(G.level_dinv(l)[:, None] * A)))
        assert abs(G.estimate_lmax(l, its=400) - lam) <= 1e-6 * lam
        assert abs(G.estimate_lmax(l, its=30) - lam) <= 0.05 * lam
