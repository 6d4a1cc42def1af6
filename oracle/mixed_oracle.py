"""CPU restatement of the reference's Taylor-Hood mixed system (Stokes).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the
product.  Pure-Python loops over cells for small grids, on top of the C
oracle's pinned helpers (quadrature, Lagrange basis, 1D cells, node points,
scalar element matrices).

Reference (M/ = Core/src/main/java/):
- numbering: TaylorHoodSpace.assembleFunctions (M/mixed/TaylorHoodSpace.java:41-92)
  puts every shape function into a TreeSet ordered by QkQkFunction.compareTo
  (M/mixed/QkQkFunction.java:82-93): velocity before pressure, velocity by
  component (SingleComponentVectorShapeFunction.compareTo,
  M/basic/SingleComponentVectorShapeFunction.java:168-173), then by node point
  under CoordinateComparator.comp (M/linalg/CoordinateComparator.java:15-27,
  ContinuousTPShapeFunction.compareTo :232-242).  `th_numbering` sorts with
  that comparator itself, not with a closed form.
- matrix: MatrixFESpace.writeCellIntegralsToMatrix (M/basic/MatrixFESpace.java:22-42)
  with MixedCellIntegral.evaluateCellIntegral (M/mixed/MixedCellIntegral.java:93-124):
  velocity GRAD_GRAD (TPVectorCellIntegral :80-88; the Frobenius product of two
  single-component gradients is the scalar product for equal components and
  exactly 0 otherwise) and DIV_VALUE (MixedTPCellIntegral :37-59) at both
  (velocity, pressure) and (pressure, velocity); element values that are
  exactly 0 are skipped; a new SparseMatrix holds a dummy (0,0) entry.
- boundary data and the pressure pin: AcceptsMatrixBoundaryValues.overWriteValue
  (M/basic/AcceptsMatrixBoundaryValues.java:25-44), whose set(1, i, i) creates
  the diagonal entry of a row that has none.
"""
from __future__ import annotations

import functools
import itertools

import numpy as np

from . import oracle as O

TOL = 1e-10   # PerformanceArguments.doubleTolerance


def coordinate_compare(a, b):
    """CoordinateComparator.comp (M/linalg/CoordinateComparator.java:15-27)"""
    for x, y in zip(a, b):
        if x > y + TOL:
            return 1
        if x < y - TOL:
            return -1
    return 0


def th_numbering(cells, kp, start, end):
    """(vel, pres): vel[c][dof] = mixed index of velocity component c at scalar Q(kp+1)
    DOF `dof`; pres[dof] = mixed index of scalar Qkp DOF `dof`."""
    d = len(cells)
    xv = O.node_coords(cells, kp + 1, start, end)
    xp = O.node_coords(cells, kp, start, end)
    funcs = [(0, c, i) for c in range(d) for i in range(len(xv))] + [(1, 0, i) for i in range(len(xp))]

    def cmp(f, g):
        # QkQkFunction.compareTo: pressure vs velocity, then component, then node point
        if f[0] != g[0]:
            return 1 if f[0] > g[0] else -1
        if f[0] == 0 and f[1] != g[1]:
            return 1 if f[1] > g[1] else -1
        pts = xv if f[0] == 0 else xp
        return coordinate_compare(pts[f[2]], pts[g[2]])

    order = sorted(funcs, key=functools.cmp_to_key(cmp))
    vel = np.zeros((d, len(xv)), dtype=np.int64)
    pres = np.zeros(len(xp), dtype=np.int64)
    for idx, (t, c, i) in enumerate(order):
        if t == 0:
            vel[c, i] = idx
        else:
            pres[i] = idx
    return vel, pres


def _local_indices(k, d):
    """local multi-indices of a cell, axis 0 fastest"""
    return [tuple((i // (k + 1) ** a) % (k + 1) for a in range(d)) for i in range((k + 1) ** d)]


def div_value_table(s, e, kv, kp, weight=1.0, rule=5):
    """B[c][iv][ip] on the cell prod [s_a, e_a]: integrateNonTensorProduct
    (M/tensorproduct/TPCellIntegral.java:78-102) of divergence * value * weight,
    divergence = component c of fastGradientInCell (ContinuousTPShapeFunction.java:173-198:
    product over axes j of derivative (j == c) or value, in axis order), value =
    fastValueInCell (:154-171); then times each axis' weight; points with the last axis fastest."""
    d = len(s)
    qp, qw = O.quad(rule)
    Q = len(qp)
    pts, wts = [], []
    for a in range(d):
        ln = e[a] - s[a]
        pts.append([qp[q] * ln + s[a] for q in range(Q)])   # Cell1D.distributeQuadrature
        wts.append([qw[q] * ln for q in range(Q)])
    lv, lp = _local_indices(kv, d), _local_indices(kp, d)

    def val1(a, l, k, x):
        ln = e[a] - s[a]
        return O.lagrange_value((x - s[a]) / ln, l, k)

    def der1(a, l, k, x):
        ln = e[a] - s[a]
        return O.lagrange_deriv((x - s[a]) / ln, l, k) / ln

    out = np.zeros((d, len(lv), len(lp)))
    for c in range(d):
        for iv, mv in enumerate(lv):
            for ip, mp in enumerate(lp):
                ret = 0.0
                for qa in itertools.product(range(Q), repeat=d):   # last axis fastest
                    x = [pts[a][qa[a]] for a in range(d)]
                    div = 1.0
                    for j in range(d):
                        div *= der1(j, mv[j], kv, x[j]) if j == c else val1(j, mv[j], kv, x[j])
                    pv = 1.0
                    for a in range(d):
                        pv *= val1(a, mp[a], kp, x[a])
                    v = div * pv * weight
                    for a in range(d):
                        v *= wts[a][qa[a]]
                    ret += v
                out[c, iv, ip] = ret
    return out


def value_grad_table(s, e, kv, kp, weight=1.0, rule=5):
    """VALUE_GRAD (MixedTPCellIntegral.java:50-57): (velocity value . pressure gradient) * weight;
    for component c the velocity value vector has one nonzero entry, so the inner product is the
    velocity function's value times the pressure gradient's component c (the other terms are
    exact zeros).  Table [c][iv][ip] like div_value_table."""
    d = len(s)
    qp, qw = O.quad(rule)
    Q = len(qp)
    pts, wts = [], []
    for a in range(d):
        ln = e[a] - s[a]
        pts.append([qp[q] * ln + s[a] for q in range(Q)])
        wts.append([qw[q] * ln for q in range(Q)])
    lv, lp = _local_indices(kv, d), _local_indices(kp, d)

    def val1(a, l, k, x):
        return O.lagrange_value((x - s[a]) / (e[a] - s[a]), l, k)

    def der1(a, l, k, x):
        ln = e[a] - s[a]
        return O.lagrange_deriv((x - s[a]) / ln, l, k) / ln

    out = np.zeros((d, len(lv), len(lp)))
    for c in range(d):
        for iv, mv in enumerate(lv):
            for ip, mp in enumerate(lp):
                ret = 0.0
                for qa in itertools.product(range(Q), repeat=d):
                    x = [pts[a][qa[a]] for a in range(d)]
                    vv = 1.0
                    for a in range(d):
                        vv *= val1(a, mv[a], kv, x[a])
                    g = 1.0
                    for j in range(d):
                        g *= der1(j, mp[j], kp, x[j]) if j == c else val1(j, mp[j], kp, x[j])
                    v = vv * g * weight
                    for a in range(d):
                        v *= wts[a][qa[a]]
                    ret += v
                out[c, iv, ip] = ret
    return out


def th_assemble(cells, kp, start, end, nu=None, div_weight=None, rule=5, vg_weight=None):
    """the mixed matrix as a dict {(row, col): value} in the mixed numbering (pattern = keys)"""
    cells = [int(c) for c in cells]
    d = len(cells)
    kv = kp + 1
    vel, pres = th_numbering(cells, kp, start, end)
    _, dv = O.dof_table(cells, kv)
    _, dp = O.dof_table(cells, kp)
    s1 = [O.cells1d(start[a], end[a], cells[a]) for a in range(d)]
    entries = {(0, 0): 0.0}
    for ci, cc in enumerate(itertools.product(*[range(n) for n in cells])):   # last axis fastest
        s = [s1[a][0][cc[a]] for a in range(d)]
        e = [s1[a][1][cc[a]] for a in range(d)]
        loc = []   # (kind, component, local index, mixed index)
        for c in range(d):
            loc += [(0, c, i, int(vel[c, dv[ci, i]])) for i in range(dv.shape[1])]
        loc += [(1, 0, i, int(pres[dp[ci, i]])) for i in range(dp.shape[1])]
        K = None
        if nu is not None:
            K = O.element_matrix(cells, kv, start, end, ci, [O.GRAD_GRAD], [nu], rule)
        B = div_value_table(s, e, kv, kp, div_weight, rule) if div_weight is not None else None
        if vg_weight is not None:   # the integrals' sum in list order: (0 + DIV_VALUE) + VALUE_GRAD
            G = value_grad_table(s, e, kv, kp, vg_weight, rule)
            B = G if B is None else B + G
        for (tv, cv, iv, gv) in loc:          # test function (row)
            for (tu, cu, iu, gu) in loc:      # trial function (column)
                val = 0.0
                if tv == 0 and tu == 0:
                    if K is not None and cv == cu:
                        val = K[iv, iu]
                elif tv != tu and B is not None:
                    if tv == 0:
                        val = B[cv, iv, iu]
                    else:
                        val = B[cu, iu, iv]
                if val != 0:
                    entries[(gv, gu)] = entries.get((gv, gu), 0.0) + val
    return entries


def to_csr(entries, n):
    keys = sorted(entries)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    cols = np.array([k[1] for k in keys], dtype=np.int32)
    vals = np.array([entries[k] for k in keys])
    for (r, _c) in keys:
        rowptr[r + 1] += 1
    rowptr = np.cumsum(rowptr)
    return O.CSR(n, rowptr, cols, vals)


def rhs_trig(cells, k, start, end, kw, amp, cos, rule=5):
    """TPRightHandSideIntegral VALUE of f = sum_j amp_j prod_a trig(kw_ja pi x_a) on a scalar
    Qk space (first-appearance DOF order), pure Python"""
    cells = [int(c) for c in cells]
    d = len(cells)
    n, dof = O.dof_table(cells, k)
    qp, qw = O.quad(rule)
    Q = len(qp)
    s1 = [O.cells1d(start[a], end[a], cells[a]) for a in range(d)]
    lv = _local_indices(k, d)
    rhs = np.zeros(n)
    for ci, cc in enumerate(itertools.product(*[range(m) for m in cells])):
        s = [s1[a][0][cc[a]] for a in range(d)]
        e = [s1[a][1][cc[a]] for a in range(d)]
        for iv, mv in enumerate(lv):
            ret = 0.0
            for qa in itertools.product(range(Q), repeat=d):
                x = [qp[qa[a]] * (e[a] - s[a]) + s[a] for a in range(d)]
                f = 0.0
                for j in range(len(amp)):
                    t = amp[j]
                    for a in range(d):
                        g = np.cos if cos[j][a] else np.sin
                        t = t * float(g(kw[j][a] * np.pi * x[a]))
                    f += t
                phi = 1.0
                for a in range(d):
                    phi *= O.lagrange_value((x[a] - s[a]) / (e[a] - s[a]), mv[a], k)
                v = phi * f
                for a in range(d):
                    v *= qw[qa[a]] * (e[a] - s[a])
                ret += v
            rhs[dof[ci, iv]] += ret
    return rhs


def overwrite_values(entries, rhs, nodes, values):
    """overWriteValue for each node in turn (AcceptsMatrixBoundaryValues.java:25-44):
    rhs_i -= A_i,node * value for every row, then column and row zeroed (entries kept),
    A_node,node = 1 (created if absent), rhs_node = value"""
    cols = {}
    for (r, c) in entries:
        cols.setdefault(c, []).append(r)
    rows = {}
    for (r, c) in entries:
        rows.setdefault(r, []).append(c)
    for node, g in zip(nodes, values):
        for r in sorted(cols.get(node, [])):
            rhs[r] += -entries[(r, node)] * g
            entries[(r, node)] = 0.0
        for c in rows.get(node, []):
            entries[(node, c)] = 0.0
        if (node, node) not in entries:
            rows.setdefault(node, []).append(node)
            cols.setdefault(node, []).append(node)
        entries[(node, node)] = 1.0
        rhs[node] = g
