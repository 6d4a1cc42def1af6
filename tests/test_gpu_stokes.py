"""Taylor-Hood Stokes path on the GPU against the mixed oracle: the assembled
saddle-point matrix (pattern bit-exact, values within 1e-12 of the largest
entry), the trigonometric body force, velocity Dirichlet data and the pressure
pin, and the block-preconditioned FGMRES solve (against scipy's direct solve
of the exported system, and against the exact solution of the reference's
StokesMG test)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import mixed_oracle as MO

pytestmark = pytest.mark.gpu


def _th(jf, cells, kp, start, end):
    s = jf.TaylorHoodSpace(start, end, cells)
    s.assembleCells()
    s.assembleFunctions(kp)
    return s


def _integrals(jf, nu, wd, rule=5, wg=None):
    out = []
    if nu is not None:
        out.append(jf.MixedCellIntegral.fromVelocityIntegral(
            jf.TPVectorCellIntegral(nu, jf.TPVectorCellIntegral.GRAD_GRAD, quadratureRule1D=rule)))
    if wd is not None:
        out.append(jf.MixedTPCellIntegral(wd, jf.MixedTPCellIntegral.DIV_VALUE, quadratureRule1D=rule))
    if wg is not None:
        out.append(jf.MixedTPCellIntegral(wg, jf.MixedTPCellIntegral.VALUE_GRAD, quadratureRule1D=rule))
    return out


def _compare(A, E, n):
    ref = MO.to_csr(E, n)
    rp, cols, vals = A.to_csr()
    assert np.array_equal(rp, ref.rowptr), "row lengths differ"
    assert np.array_equal(cols, ref.cols), "pattern differs"
    scale = np.max(np.abs(ref.vals))
    err = np.max(np.abs(vals - ref.vals)) if len(vals) else 0.0
    assert err <= 1e-12 * scale, err


@pytest.mark.parametrize("cells,kp,nu,wd", [((3, 4), 1, 0.7, -1.0), ((2, 3), 1, None, 1.0), ((4, 2), 1, 2.0, None),
                                            ((2, 2), 2, 1.0, -1.0), ((1, 2, 2), 1, 1.5, -1.0)])
def test_th_assembly_matches_oracle(jf, oracle, cells, kp, nu, wd):
    d = len(cells)
    start, end = [0.5, -1.0, 0.0][:d], [2.0, 1.0, 0.75][:d]
    print(f"th assembly cells={cells} kp={kp} nu={nu} wd={wd}", flush=True)
    s = _th(jf, cells, kp, start, end)
    A = s.writeCellIntegralsToMatrix(_integrals(jf, nu, wd))
    E = MO.th_assemble(cells, kp, start, end, nu, wd)
    assert A.getRows() == s.ndofs()
    _compare(A, E, s.ndofs())


@pytest.mark.parametrize("cells,kp,nu,wd,wg", [((3, 4), 1, None, None, 1.0), ((3, 2), 1, 0.7, -1.0, 0.5),
                                               ((2, 2), 2, None, 1.0, -2.0), ((1, 2, 2), 1, 1.5, None, -1.0)])
def test_th_assembly_value_grad(jf, oracle, cells, kp, nu, wd, wg):
    """MixedTPCellIntegral VALUE_GRAD alone and summed with DIV_VALUE in one element value"""
    d = len(cells)
    start, end = [0.5, -1.0, 0.0][:d], [2.0, 1.0, 0.75][:d]
    print(f"th assembly value_grad cells={cells} kp={kp} nu={nu} wd={wd} wg={wg}", flush=True)
    s = _th(jf, cells, kp, start, end)
    A = s.writeCellIntegralsToMatrix(_integrals(jf, nu, wd, wg=wg))
    E = MO.th_assemble(cells, kp, start, end, nu, wd, vg_weight=wg)
    _compare(A, E, s.ndofs())


@pytest.mark.parametrize("rule", [3, 4])
def test_th_assembly_other_quadrature_rules(jf, oracle, rule):
    """Gauss3 and Gauss4 instead of the default Gauss5"""
    cells, kp = (3, 2), 1
    start, end = [0.0, 0.5], [1.0, 1.25]
    print(f"th assembly rule={rule}", flush=True)
    s = _th(jf, cells, kp, start, end)
    A = s.writeCellIntegralsToMatrix(_integrals(jf, 1.3, -1.0, rule))
    E = MO.th_assemble(cells, kp, start, end, 1.3, -1.0, rule)
    _compare(A, E, s.ndofs())


def test_th_assembly_rejects_face_quadrature(jf):
    """Gauss-Lobatto points lie on cell faces, where the reference evaluates the mixed
    integrands in the first support cell containing the point: not modelled, refused"""
    s = _th(jf, (2, 2), 1, [0.0, 0.0], [1.0, 1.0])
    with pytest.raises(jf.UnsupportedOperationException):
        s.writeCellIntegralsToMatrix(_integrals(jf, 1.0, -1.0, 15))


def test_th_argument_errors(jf):
    s = _th(jf, (2, 2), 1, [0.0, 0.0], [1.0, 1.0])
    other = _th(jf, (3, 2), 1, [0.0, 0.0], [1.0, 1.0])
    ints = _integrals(jf, 1.0, -1.0)
    # velocity and pressure spaces on different grids
    h = jf.api.C.c_void_p()
    kinds = np.array([0, 1], dtype=np.int32)
    w = np.array([1.0, -1.0])
    rc = jf.lib().jfem_th_assemble(s.velocity.h, other.pressure.h, 2, jf._native.iptr(kinds),
                                   jf._native.dptr(w), 5, jf.api.C.byref(h))
    assert rc == 1
    # degrees that are not k+1 / k
    rc = jf.lib().jfem_th_assemble(s.pressure.h, s.pressure.h, 2, jf._native.iptr(kinds),
                                   jf._native.dptr(w), 5, jf.api.C.byref(h))
    assert rc == 1
    # unknown integral kind, two integrals of one kind
    with pytest.raises(jf.UnsupportedOperationException):
        jf.MixedTPCellIntegral(1.0, "GradValue")
    for bad in ([0, 0], [2, 2], [0, 7]):
        bad = np.array(bad, dtype=np.int32)
        rc = jf.lib().jfem_th_assemble(s.velocity.h, s.pressure.h, 2, jf._native.iptr(bad), jf._native.dptr(w), 5,
                                       jf.api.C.byref(h))
        assert rc == 3
    A = s.writeCellIntegralsToMatrix(ints)
    rhs = jf.DenseVector(s.ndofs())
    # a body force with the wrong number of components
    f = jf.VectorTrigSource([jf.TrigSource([[1, 1]], [1.0])])
    with pytest.raises(jf.IllegalArgumentException):
        s.writeCellIntegralsToRhs([jf.MixedRightHandSideIntegral.fromVelocityIntegral(
            jf.TPVectorRightHandSideIntegral(f, "Value"))], rhs)
    # a velocity hierarchy of another size
    mg, _ = _velocity_mg(jf, (4, 4), 1.0, 0)
    with pytest.raises(jf.IllegalArgumentException):
        jf.StokesBlockPreconditioner(s, mg, 1.0)


def _lid(x):
    """u = (1, 0[, 0]) on the face x_{d-1} = 1 away from its edges, 0 elsewhere on the boundary"""
    d = len(x)
    top = abs(x[d - 1] - 1.0) < 1e-12 and all(0 < x[a] < 1 for a in range(d - 1))
    return (1.0 if top else 0.0,) + (0.0,) * (d - 1)


def test_th_rhs_boundary_and_pin_match_oracle(jf, oracle):
    cells, kp = (4, 3), 1
    start, end = [0.0, 0.0], [1.0, 1.0]
    print("th rhs/boundary/pin", flush=True)
    s = _th(jf, cells, kp, start, end)
    A = s.writeCellIntegralsToMatrix(_integrals(jf, 1.0, -1.0))
    f = jf.VectorTrigSource([jf.TrigSource([[1, 2], [2, 1]], [0.5, -1.5], [[1, 0], [0, 0]]),
                             jf.TrigSource([[1, 1]], [2.0], [[0, 1]])])
    rhs = jf.DenseVector(s.ndofs())
    s.writeCellIntegralsToRhs([jf.MixedRightHandSideIntegral.fromVelocityIntegral(
        jf.TPVectorRightHandSideIntegral(f, jf.TPVectorRightHandSideIntegral.VALUE))], rhs)
    # oracle right-hand side: per component the scalar Q2 load, placed by the mixed numbering
    vel, pres = MO.th_numbering(cells, kp, start, end)
    n = s.ndofs()
    bref = np.zeros(n)
    for c, src in enumerate(f.components):
        r = MO.rhs_trig(cells, kp + 1, start, end, src.kw.tolist(), src.amp.tolist(), src.cos.tolist())
        bref[vel[c]] += r
    b = rhs.numpy()
    assert np.max(np.abs(b - bref)) <= 1e-13 * np.max(np.abs(bref))
    # velocity boundary data (lid) and the first pressure unknown pinned to 0
    s.writeVelocityBoundaryValuesTo(_lid, A, rhs)
    s.overWriteValue(s.getVelocitySize(), 0.0, A, rhs)
    E = MO.th_assemble(cells, kp, start, end, 1.0, -1.0)
    xv = oracle.node_coords(cells, kp + 1, start, end)
    bnodes = oracle.boundary_nodes(cells, kp + 1)
    nodes, vals = [], []
    for c in range(2):
        for i in bnodes:
            nodes.append(int(vel[c, i]))
            vals.append(_lid(xv[i])[c])
    order = np.argsort(nodes)
    nodes = [nodes[i] for i in order]
    vals = [vals[i] for i in order]
    MO.overwrite_values(E, bref, nodes, vals)
    MO.overwrite_values(E, bref, [s.getVelocitySize()], [0.0])
    _compare(A, E, n)
    b = rhs.numpy()
    assert np.max(np.abs(b - bref)) <= 1e-12 * np.max(np.abs(bref))


def _stokes_system(jf, cells, nu, source=None, boundary=None):
    start, end = [0.0] * len(cells), [1.0] * len(cells)
    s = _th(jf, cells, 1, start, end)
    A = s.writeCellIntegralsToMatrix(_integrals(jf, nu, -1.0))
    rhs = jf.DenseVector(s.ndofs())
    if source is not None:
        s.writeCellIntegralsToRhs([jf.MixedRightHandSideIntegral.fromVelocityIntegral(
            jf.TPVectorRightHandSideIntegral(source, jf.TPVectorRightHandSideIntegral.VALUE))], rhs)
    s.writeVelocityBoundaryValuesTo(boundary, A, rhs)
    s.overWriteValue(s.getVelocitySize(), 0.0, A, rhs)
    return s, A, rhs


def _velocity_mg(jf, cells, nu, precision, smoother=None, ctx=None):
    spaces = []
    c = list(cells)
    levels = []
    while min(c) >= 2 and len(levels) < 8:
        levels.append(list(c))
        if min(c) % 2:
            break
        c = [x // 2 for x in c]
    for lc in reversed(levels):
        sp_ = jf.ContinuousTPFESpace([0.0] * len(lc), [1.0] * len(lc), lc, ctx)
        sp_.assembleCells()
        sp_.assembleFunctions(2)
        spaces.append(sp_)
    sm = smoother or jf.ChebyshevSmoother(degree=3, runs=1)
    return jf.MGPreconditioner(spaces, [jf.TPCellIntegral(nu, jf.TPCellIntegral.GRAD_GRAD)], sm, cycles=1,
                               fmg=False, precision=precision), spaces


@pytest.mark.parametrize("precision,mass_its", [(0, 0), (1, 0), (0, 4), (1, 4)])
def test_stokes_fgmres_matches_direct_solve(jf, precision, mass_its):
    cells, nu = (16, 16), 0.5
    print(f"stokes fgmres vs spsolve precision={precision} mass_its={mass_its}", flush=True)
    src = jf.VectorTrigSource([jf.TrigSource([[1, 2]], [3.0]), jf.TrigSource([[2, 1]], [-2.0], [[1, 0]])])
    s, A, rhs = _stokes_system(jf, cells, nu, src, _lid)
    mg, _ = _velocity_mg(jf, cells, nu, precision)
    prec = jf.StokesBlockPreconditioner(s, mg, nu, mass_its=mass_its)
    x = jf.DenseVector(s.ndofs())
    info, hist = jf.solveStokesFGMRES(A, prec, rhs, x, 1e-12, restart=50, maxit=400)
    rp, cols, vals = A.to_csr()
    M = sp.csr_matrix((vals, cols, rp), shape=(s.ndofs(), s.ndofs()))
    b = rhs.numpy()
    xd = spla.spsolve(M.tocsc(), b)
    xs = x.numpy()
    assert np.linalg.norm(M @ xs - b) <= 1.01e-12 * np.linalg.norm(b)
    # Pinning one corner pressure leaves the constant-pressure mode almost null (it reaches
    # only the interior velocity nodes of the corner cell), so the iterative solution may
    # differ from the direct one by a constant in the free pressures: compare the velocity
    # directly and the pressure modulo a constant
    nv2 = s.getVelocitySize()
    assert np.linalg.norm(xs[:nv2] - xd[:nv2]) <= 1e-8 * np.linalg.norm(xd[:nv2])
    dp = xs[nv2 + 1:] - xd[nv2 + 1:]
    assert np.linalg.norm(dp - dp.mean()) <= 1e-8 * np.linalg.norm(xd[nv2:])
    assert xs[nv2] == 0.0 or abs(xs[nv2]) <= 1e-12 * np.linalg.norm(b)
    assert info.iterations < 120, info.iterations
    assert np.all(hist[1:] <= hist[:-1] * (1 + 1e-6) + 1e-14)   # GMRES residuals do not increase


@pytest.mark.parametrize("mass_its", [0, 4])
def test_stokes_fgmres_3d_matches_direct_solve(jf, mass_its):
    cells, nu = (4, 4, 4), 1.0
    print(f"3d stokes fgmres vs spsolve mass_its={mass_its}", flush=True)
    src = jf.VectorTrigSource([jf.TrigSource([[1, 2, 1]], [3.0]), jf.TrigSource([[2, 1, 1]], [-2.0], [[1, 0, 0]]),
                               jf.TrigSource([[1, 1, 2]], [1.0], [[0, 0, 1]])])
    s, A, rhs = _stokes_system(jf, cells, nu, src, _lid)
    mg, _ = _velocity_mg(jf, cells, nu, 0)
    prec = jf.StokesBlockPreconditioner(s, mg, nu, mass_its=mass_its)
    x = jf.DenseVector(s.ndofs())
    info, _ = jf.solveStokesFGMRES(A, prec, rhs, x, 1e-12, restart=50, maxit=400)
    rp, cols, vals = A.to_csr()
    M = sp.csr_matrix((vals, cols, rp), shape=(s.ndofs(), s.ndofs()))
    b = rhs.numpy()
    xd = spla.spsolve(M.tocsc(), b)
    xs = x.numpy()
    assert np.linalg.norm(M @ xs - b) <= 1.01e-12 * np.linalg.norm(b)
    nv3 = s.getVelocitySize()
    assert np.linalg.norm(xs[:nv3] - xd[:nv3]) <= 1e-8 * np.linalg.norm(xd[:nv3])
    dp = xs[nv3 + 1:] - xd[nv3 + 1:]
    assert np.linalg.norm(dp - dp.mean()) <= 1e-8 * np.linalg.norm(xd[nv3:])


@pytest.mark.parametrize("precision", [0, 1])
def test_stokes_concurrent_component_cycles(jf, precision):
    """component 1's V-cycle on a second context (stream), joined by events: the same
    solution bit for bit as the sequential preconditioner"""
    cells, nu = (32, 32), 1.0
    print(f"stokes concurrent components precision={precision}", flush=True)
    src = jf.VectorTrigSource([jf.TrigSource([[1, 2]], [3.0]), jf.TrigSource([[2, 1]], [-2.0], [[1, 0]])])
    s, A, rhs = _stokes_system(jf, cells, nu, src, _lid)
    mg, _ = _velocity_mg(jf, cells, nu, precision)
    ctx2 = jf.Context(0)
    mg2, _ = _velocity_mg(jf, cells, nu, precision, ctx=ctx2)
    xs = []
    for comps in ((), (mg2,)):
        prec = jf.StokesBlockPreconditioner(s, mg, nu, mass_its=3, component_mgs=comps)
        x = jf.DenseVector(s.ndofs())
        info, _ = jf.solveStokesFGMRES(A, prec, rhs, x, 1e-10, restart=50, maxit=300)
        assert info.resnorm <= 1.01e-10 * rhs.euclidianNorm()
        xs.append(x.numpy())
    assert np.array_equal(xs[0], xs[1])


def _reference_source(jf):
    """StokesReferenceSolution.rightHandSide (Core/src/test/java/examples/StokesReferenceSolution.java:14-37)
    as sums of sine/cosine products:
    f_x = -pi^2 sin(2 pi y) + 2 pi^2 cos(2 pi x) sin(2 pi y) + pi cos(pi x) sin(pi y)
    f_y =  pi^2 sin(2 pi x) - 2 pi^2 sin(2 pi x) cos(2 pi y) + pi sin(pi x) cos(pi y)"""
    pi = np.pi
    fx = jf.TrigSource([[0, 2], [2, 2], [1, 1]], [-pi ** 2, 2 * pi ** 2, pi], [[1, 0], [1, 0], [1, 0]])
    fy = jf.TrigSource([[2, 0], [2, 2], [1, 1]], [pi ** 2, -2 * pi ** 2, pi], [[0, 1], [0, 1], [0, 1]])
    return jf.VectorTrigSource([fx, fy])


def _reference_velocity(p):
    x, y = p[..., 0], p[..., 1]
    s, c = np.sin, np.cos
    return np.stack([-s(np.pi * x) ** 2 * c(np.pi * y) * s(np.pi * y),
                     s(np.pi * x) * c(np.pi * x) * s(np.pi * y) ** 2], axis=-1)


def test_stokes_reference_solution_convergence(jf):
    """StokesMG.testConvergenceBoundary (Core/src/test/java/examples/StokesMG.java:26-159):
    4 * 2^r cells, zero velocity on the boundary, first pressure pinned to 0, error of
    the velocity over generatePlotPoints(20) below 0.2 * 0.5^(2r); solved here with the
    block-preconditioned FGMRES instead of the reference's monolithic multigrid"""
    errs = []
    for r in range(4):
        cells = (4 * 2 ** r, 4 * 2 ** r)
        print(f"stokes convergence r={r}", flush=True)
        s, A, rhs = _stokes_system(jf, cells, 1.0, _reference_source(jf), None)
        mg, _ = _velocity_mg(jf, cells, 1.0, 0)
        prec = jf.StokesBlockPreconditioner(s, mg, 1.0)
        x = jf.DenseVector(s.ndofs())
        info, _ = jf.solveStokesFGMRES(A, prec, rhs, x, 1e-10, restart=50, maxit=400)
        assert info.resnorm <= 1.01e-10 * rhs.euclidianNorm()
        g = np.linspace(0, 1, 20)
        pts = np.array([[a, b] for a in g for b in g])   # CartesianGrid.generatePlotPoints(20)
        uh = np.stack([jf.ScalarFESpaceFunction(s.velocity, jf.DenseVector(s.velocity_component(x, c))).value(pts)
                       for c in range(2)], axis=-1)
        err = float(np.sqrt(np.sum(np.sum((uh - _reference_velocity(pts)) ** 2, axis=-1))))
        errs.append(err)
        assert err < 0.2 * 0.5 ** (2 * r), (r, err)
    # Q2 velocity: the pointwise error falls like h^3
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > 2.5, rates


def test_stokes_learned_velocity_block(jf):
    """The Taylor-Hood preconditioner with the learned smoother on the finest level of the FP32
    velocity multigrid (network 2 + dim input channels, trained with the two-grid loss on the
    scalar velocity hierarchy of the same grid): FGMRES meets the tolerance on the true residual,
    the velocity matches scipy's direct solve and the pressure does modulo a constant, and the
    network never falls back to Chebyshev.  The iteration counts of the learned and the
    Chebyshev velocity blocks are printed."""
    cells, nu = (32, 32), 1.0
    print("stokes learned velocity block", flush=True)
    src = jf.VectorTrigSource([jf.TrigSource([[1, 2]], [3.0]), jf.TrigSource([[2, 1]], [-2.0], [[1, 0]])])
    s, A, rhs = _stokes_system(jf, cells, nu, src, _lid)
    ncell = cells[0] * cells[1]
    mgt, _ = _velocity_mg(jf, cells, nu, 0)
    net = jf.SmootherNet(2, 4, seed=7)
    mgt.set_learned(net, np.zeros(ncell), runs=1, fallback_ratio=0.5)
    h = mgt.train(rhs_per_field=120, cycles_per_rhs=8, lr=1e-4, seed=1, mode=5)
    assert np.all(np.isfinite(h)) and np.mean(h[-16:]) < 0.5 * np.mean(h[:8]), (np.mean(h[:8]), np.mean(h[-16:]))
    mg, _ = _velocity_mg(jf, cells, nu, 1)
    mg.set_learned(net, np.zeros(ncell), runs=1, fallback_ratio=0.5)
    rp, cols, vals = A.to_csr()
    M = sp.csr_matrix((vals, cols, rp), shape=(s.ndofs(), s.ndofs()))
    b = rhs.numpy()
    xd = spla.spsolve(M.tocsc(), b)
    nv2 = s.getVelocitySize()
    its = {}
    for learned in (True, False):
        mg.set_learned_active(learned)
        prec = jf.StokesBlockPreconditioner(s, mg, nu, mass_its=3)
        x = jf.DenseVector(s.ndofs())
        info, _ = jf.solveStokesFGMRES(A, prec, rhs, x, 1e-12, restart=50, maxit=400)
        xs = x.numpy()
        assert np.linalg.norm(M @ xs - b) <= 1.01e-12 * np.linalg.norm(b), learned
        assert np.linalg.norm(xs[:nv2] - xd[:nv2]) <= 1e-8 * np.linalg.norm(xd[:nv2]), learned
        dp = xs[nv2 + 1:] - xd[nv2 + 1:]
        assert np.linalg.norm(dp - dp.mean()) <= 1e-8 * np.linalg.norm(xd[nv2:]), learned
        assert info.iterations < 120, (learned, info.iterations)
        its[learned] = info.iterations
        if learned:
            assert not mg.learned_status()[1], "the learned smoother fell back"
    print(f"FGMRES iterations to 1e-12: learned velocity block {its[True]}, Chebyshev {its[False]}")
