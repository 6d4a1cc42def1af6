"""Pins the Taylor-Hood (Stokes) restatement in oracle/mixed_oracle.py against
the reference's golden DIV_VALUE vector, the TreeSet order replayed with the
reference's comparator, exact integrals of polynomial fields and the C
oracle's pinned elimination.  CPU only."""

import numpy as np
import pytest

from oracle import mixed_oracle as MO
from oracle import oracle as O

# MixedCellIntegralTest.testMixedIntegral (Core/src/test/java/mixed/MixedCellIntegralTest.java:72-96):
# cell [3,5]x[0,4], Q1 velocity function i (component i // 4) against the first Q1
# pressure function, DIV_VALUE with the default weight, checked there to 1e-2
VALUES_DV = [-0.666, 0.666, -0.333, 0.333, -0.333, -0.1666, 0.333, 0.1666]


def test_div_value_golden(oracle):
    B = MO.div_value_table([3.0, 0.0], [5.0, 4.0], 1, 1, 1.0)
    got = [B[i // 4, i % 4, 0] for i in range(8)]
    assert np.allclose(got, VALUES_DV, atol=1e-2)
    # the same values exactly: (int L_a' L_0 dx)(int L_b L_0 dy) on a 2 x 4 cell
    exact = [-2 / 3, 2 / 3, -1 / 3, 1 / 3, -1 / 3, -1 / 6, 1 / 3, 1 / 6]
    assert np.allclose(got, exact, rtol=0, atol=1e-14)


VALUES_VG = [-0.666, -0.666, -0.333, -0.333, -0.333, -0.1666, -0.333, -0.1666]


def test_value_grad_golden(oracle):
    """MixedCellIntegralTest.testMixedIntegral's VALUE_GRAD vector (:84, :92), to 1e-2 as there and
    exactly: (int L_a L_0' dx)(int L_b L_0 dy) and its y counterpart on the 2 x 4 cell"""
    G = MO.value_grad_table([3.0, 0.0], [5.0, 4.0], 1, 1, 1.0)
    got = [G[i // 4, i % 4, 0] for i in range(8)]
    assert np.allclose(got, VALUES_VG, atol=1e-2)
    exact = [-2 / 3, -2 / 3, -1 / 3, -1 / 3, -1 / 3, -1 / 6, -1 / 3, -1 / 6]
    assert np.allclose(got, exact, rtol=0, atol=1e-14)
    # integration by parts on the cell: VALUE_GRAD + DIV_VALUE = boundary term, which for the
    # pressure function 0 and velocity function 3 (nodes (5,4), pressure node (3,0)) vanishes
    B = MO.div_value_table([3.0, 0.0], [5.0, 4.0], 1, 1, 1.0)
    assert abs(G[0, 3, 0] + B[0, 3, 0]) < 1e-14


def test_value_grad_is_minus_div_value_for_interior_velocity(oracle):
    """integration by parts over the domain: for a velocity function vanishing on the boundary,
    the assembled VALUE_GRAD row equals minus the DIV_VALUE row (boundary rows do not)"""
    cells, kp = (3, 2), 1
    start, end = [0.0, -1.0], [1.5, 1.0]
    G = MO.th_assemble(cells, kp, start, end, vg_weight=1.0)
    B = MO.th_assemble(cells, kp, start, end, div_weight=1.0)
    vel, _ = MO.th_numbering(cells, kp, start, end)
    xv = O.node_coords(cells, kp + 1, start, end)
    inner = {int(vel[c, i]) for c in range(2) for i in range(len(xv))
             if all(start[a] + 1e-12 < xv[i][a] < end[a] - 1e-12 for a in range(2))}
    nv = 2 * len(xv)
    checked = broken = 0
    for r in range(nv):
        row_g = {c: v for (rr, c), v in G.items() if rr == r and c >= nv}
        row_b = {c: v for (rr, c), v in B.items() if rr == r and c >= nv}
        diff = max(abs(row_g.get(c, 0.0) + row_b.get(c, 0.0)) for c in set(row_g) | set(row_b))
        if r in inner:
            assert diff < 1e-13, (r, diff)
            checked += 1
        elif diff > 1e-3:
            broken += 1
    assert checked > 0 and broken > 0


def test_div_value_golden_detects_swapped_component(oracle):
    B = MO.div_value_table([3.0, 0.0], [5.0, 4.0], 1, 1, 1.0)
    swapped = [B[1 - i // 4, i % 4, 0] for i in range(8)]
    assert not np.allclose(swapped, VALUES_DV, atol=1e-2)


def _lex(N, p):
    r = 0
    for a in range(len(N)):
        r = r * N[a] + p[a]
    return r


@pytest.mark.parametrize("cells,kp", [((3, 2), 1), ((2, 1, 2), 1), ((2, 3), 2)])
def test_th_numbering_is_component_major_lexicographic(oracle, cells, kp):
    """the comparator replay equals velocity components first, then pressure, each in
    lexicographic lattice order with the last axis fastest (the closed form of the HIP path)"""
    O = oracle
    d = len(cells)
    start, end = [0.25] * d, [1.5 + a for a in range(d)]
    vel, pres = MO.th_numbering(cells, kp, start, end)
    NV = [(kp + 1) * n + 1 for n in cells]
    NP = [kp * n + 1 for n in cells]
    nv, npr = int(np.prod(NV)), int(np.prod(NP))
    xv = O.node_coords(cells, kp + 1, start, end)
    xp = O.node_coords(cells, kp, start, end)
    h = [(end[a] - start[a]) / cells[a] for a in range(d)]
    for c in range(d):
        for dof in range(nv):
            p = [int(round((xv[dof, a] - start[a]) / h[a] * (kp + 1))) for a in range(d)]
            assert vel[c, dof] == c * nv + _lex(NV, p)
    for dof in range(npr):
        p = [int(round((xp[dof, a] - start[a]) / h[a] * kp)) for a in range(d)]
        assert pres[dof] == d * nv + _lex(NP, p)


def _gauss_2d(f, s, e, n=8):
    x, w = np.polynomial.legendre.leggauss(n)
    tot = 0.0
    for i in range(n):
        for j in range(n):
            px = s[0] + (x[i] + 1) / 2 * (e[0] - s[0])
            py = s[1] + (x[j] + 1) / 2 * (e[1] - s[1])
            tot += w[i] * w[j] * f(px, py)
    return tot * (e[0] - s[0]) * (e[1] - s[1]) / 4


def test_th_assemble_exact_integrals(oracle):
    """for u in Q2^2 and p in Q1 interpolated: p^T B u = w int div(u) p and
    u^T A u = nu int |grad u|^2 (independent Gauss-Legendre integration);
    no entries couple the two velocity components"""
    O = oracle
    cells, kp = (3, 2), 1
    start, end = [0.5, -1.0], [2.0, 1.0]
    nu, wd = 0.7, -1.0
    E = MO.th_assemble(cells, kp, start, end, nu, wd)
    vel, pres = MO.th_numbering(cells, kp, start, end)
    xv = O.node_coords(cells, 2, start, end)
    xp = O.node_coords(cells, 1, start, end)
    nv = xv.shape[0]
    n = 2 * nv + xp.shape[0]
    u0 = lambda x, y: x * x * y - x * y + 0.3 * y * y
    u1 = lambda x, y: y * y + 3 * x * y - x * x * y * y
    p = lambda x, y: 1 + 2 * x - y + x * y
    U = np.zeros(n)
    P = np.zeros(n)
    U[vel[0]] = u0(xv[:, 0], xv[:, 1])
    U[vel[1]] = u1(xv[:, 0], xv[:, 1])
    P[pres] = p(xp[:, 0], xp[:, 1])
    pBu = sum(v * P[r] * U[c] for (r, c), v in E.items())
    uBp = sum(v * U[r] * P[c] for (r, c), v in E.items())
    div = lambda x, y: (2 * x * y - y) + (2 * y + 3 * x - 2 * x * x * y)
    exact = wd * _gauss_2d(lambda x, y: div(x, y) * p(x, y), start, end)
    assert abs(pBu - exact) <= 1e-12 * max(1, abs(exact))
    assert abs(uBp - exact) <= 1e-12 * max(1, abs(exact))
    uAu = sum(v * U[r] * U[c] for (r, c), v in E.items())
    g2 = lambda x, y: ((2 * x * y - y) ** 2 + (x * x - x + 0.6 * y) ** 2 + (3 * y - 2 * x * y * y) ** 2
                       + (2 * y + 3 * x - 2 * x * x * y) ** 2)
    assert abs(uAu - nu * _gauss_2d(g2, start, end)) <= 1e-11 * nu * _gauss_2d(g2, start, end)
    for (r, c) in E:
        if r < 2 * nv and c < 2 * nv and (r, c) != (0, 0):
            assert (r < nv) == (c < nv)
        assert E.get((c, r)) == E[(r, c)]   # symmetric, bit for bit
    assert (0, 0) in E


def test_rhs_trig_matches_c_oracle_and_integrates_exactly(oracle):
    O = oracle
    cells, k = (3, 4), 2
    start, end = [0.0, 0.0], [1.0, 1.0]
    kw = [[1, 2], [3, 1]]
    amp = [0.7, -1.1]
    r1 = MO.rhs_trig(cells, k, start, end, kw, amp, [[0, 0], [0, 0]])
    r2 = O.rhs_sines(cells, k, start, end, kw, amp)
    assert np.max(np.abs(r1 - r2)) <= 1e-15 * np.max(np.abs(r2))
    # partition of unity: sum_i rhs_i = int f; int cos(2 pi x) sin(pi y) = 0, int cos(0) sin(2 pi y) = 0,
    # int cos(pi x) sin(pi y) dx dy = 0, int sin(pi x) sin(pi y) = 4 / pi^2
    kw = [[2, 1], [0, 2], [1, 1], [1, 1]]
    cos = [[1, 0], [1, 0], [1, 0], [0, 0]]
    amp = [1.0, 2.0, 3.0, 0.5]
    r = MO.rhs_trig((8, 8), k, start, end, kw, amp, cos)
    assert abs(r.sum() - 0.5 * 4 / np.pi ** 2) < 1e-9
    # a wrong factor (sin for cos) changes the integral: int sin(2 pi x) sin(pi y) = 0 but
    # int sin(0 x) = 0 and int sin(pi x) sin(pi y) = 4/pi^2
    rw = MO.rhs_trig((8, 8), k, start, end, kw, amp, [[0, 0], [1, 0], [0, 0], [0, 0]])
    assert abs(rw.sum() - 0.5 * 4 / np.pi ** 2) > 1e-3


def test_overwrite_values_matches_c_oracle(oracle):
    O = oracle
    cells, k = (3, 3), 2
    A = O.assemble(cells, k, [0, 0], [1, 1], [O.GRAD_GRAD], [1.3])
    E = {}
    for r in range(A.n):
        for p in range(A.rowptr[r], A.rowptr[r + 1]):
            E[(r, int(A.cols[p]))] = A.vals[p]
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.n)
    nodes = O.boundary_nodes(cells, k)
    g = rng.standard_normal(len(nodes))
    b1 = b.copy()
    O.dirichlet(A, b1, nodes, g)
    b2 = b.copy()
    MO.overwrite_values(E, b2, list(nodes), list(g))
    assert np.allclose(b1, b2, rtol=1e-13, atol=1e-13)
    for r in range(A.n):
        for p in range(A.rowptr[r], A.rowptr[r + 1]):
            assert E[(r, int(A.cols[p]))] == A.vals[p]


def test_overwrite_values_creates_missing_diagonal():
    E = {(0, 0): 2.0, (0, 1): 1.0, (1, 0): 1.0}
    rhs = np.array([1.0, 5.0])
    MO.overwrite_values(E, rhs, [1], [3.0])
    assert E == {(0, 0): 2.0, (0, 1): 0.0, (1, 0): 0.0, (1, 1): 1.0}
    assert np.allclose(rhs, [1.0 - 3.0, 3.0])


def test_reference_source_expansion():
    """the sine/cosine expansion used for StokesReferenceSolution.rightHandSide
    (Core/src/test/java/examples/StokesReferenceSolution.java:14-37) in the GPU tests
    equals the reference's expression"""
    import jfem_amd as jf
    from tests.test_gpu_stokes import _reference_source
    pi = np.pi
    pts = np.random.default_rng(0).random((50, 2))
    x, y = pts[:, 0], pts[:, 1]
    s, c = np.sin, np.cos
    fx = (2 * pi * pi * c(pi * x) * c(pi * x) * s(pi * y) * c(pi * y)
          - 6 * pi * pi * s(pi * x) * s(pi * x) * s(pi * y) * c(pi * y) + pi * c(pi * x) * s(pi * y))
    fy = (-2 * pi * pi * c(pi * y) * c(pi * y) * s(pi * x) * c(pi * x)
          + 6 * pi * pi * s(pi * y) * s(pi * y) * s(pi * x) * c(pi * x) + s(pi * x) * pi * c(pi * y))
    f = _reference_source(jf)(pts)
    assert np.allclose(f[:, 0], fx, atol=1e-12) and np.allclose(f[:, 1], fy, atol=1e-12)
