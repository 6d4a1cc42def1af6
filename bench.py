"""Benchmark of the JFEM Qk solve path on one MI355X.

Workload (N=1): 3D Poisson -div grad u = f on the unit cube, Q2 ContinuousTPFESpace
on 96^3 cells (7,189,057 DOFs), f = seeded sum of 16 3D sinusoids, zero Dirichlet,
geometric multigrid on levels 3..96 cells (coarsest 3^3 solved directly), to
||r|| <= 1e-8 ||b||:
  - FP64 MG-PCG: V(2,2) with Chebyshev-Jacobi degree 4 preconditioning CG (the
    plain drop-in path);
  - FP32 MG inside FP64 flexible CG, same smoother;
  - learned + FP32 MG inside FP64 flexible CG: the finest-level smoother is the
    BF16-MFMA network, trained here on the same operator on 24^3 cells (one
    application per smoothing call, Chebyshev fallback).
`value` is the learned + FP32 time (the metric compares it with the FP64 MG-PCG
time, also in the line); with --no-fp32 / --no-learned it is the FP64 time.

A step is one solve from x0 = 0 with the matrices, hierarchy and right-hand side
already resident in HBM.  The line also reports assembly time, the FP64 SpMV's
roofline (with the measured HBM traffic and a streaming-read calibration), and
the CPU oracle timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_problem(jf, base, nlev, degree, smoother, precision, seed, nterms, cell_coeff=None, galerkin=False):
    d = len(base)
    spaces = []
    for l in range(nlev):
        s = jf.ContinuousTPFESpace([0.0] * d, [1.0] * d, [c << l for c in base])
        s.assembleCells()
        s.assembleFunctions(degree)
        spaces.append(s)
    ints = [jf.TPCellIntegral("GradGrad")]
    t0 = time.perf_counter()
    G = jf.MGPreconditioner(spaces, ints, smoother, cycles=1, fmg=False, precision=precision, cell_coeff=cell_coeff,
                            galerkin=galerkin)
    t_mg = time.perf_counter() - t0
    return spaces, G, t_mg


def rhs_for(jf, space, seed, nterms):
    from jfem_amd.workloads import sinusoid_terms
    kw, amp = sinusoid_terms(seed, nterms, space.getDimension())
    space.initializeRhs()
    space.writeCellIntegralsToRhs([jf.TPRightHandSideIntegral(jf.SinusoidSource(kw, amp))], space.getRhs())
    r = space.getRhs().numpy()
    r[space.boundary_nodes()] = 0.0
    return jf.DenseVector(r), r


def cpu_baseline_full(G, nlev, base, k, rhs, tol, budget_s=30.0, x_gpu=None):
    """The CPU baseline at the full workload size: the OpenMP restatement of the reference's
    CPU solve path (oracle/cpu_baseline.c: row-parallel SparseMvMul products, mgStepV with the
    same Chebyshev-Jacobi V(2,2) degree-4 smoother, solvePCG from x0 = 0 to the same absolute
    tolerance) on all host cores, over the same level systems (exported from the GPU hierarchy,
    which the tests show bit-exact with the oracle's assembly).  Solves are repeated while they
    fit in budget_s (at most 3); value = the median solve time."""
    import scipy.sparse as sps

    from oracle import cpu_baseline as CB
    t0 = time.perf_counter()
    levels = [G.getSystem(l).to_csr() for l in range(nlev)]
    t_export = time.perf_counter() - t0
    t0 = time.perf_counter()
    M = CB.CpuMG(len(base), k, base, levels, degree=4, runs=2, ratio=20.0)
    t_setup = time.perf_counter() - t0
    runs = []
    while len(runs) < 3 and (not runs or sum(r[0] for r in runs) + runs[-1][0] <= budget_s):
        x, it, secs, res = M.pcg(rhs, tol)
        runs.append((secs, it, res))
        log(f"CPU baseline solve: {it} its, {secs:.2f}s on {CB.threads()} threads")
    n = len(rhs)
    nnz = int(levels[-1][0][-1])
    spmv_s = M.spmv_seconds(5)
    out = {"value": float(np.median([r[0] for r in runs])), "unit": "s", "cores": CB.threads(), "kind": "port",
           "sample": f"full workload ({n} DOFs, {nnz} nnz): OpenMP C restatement of the reference's CPU path "
                     f"(row-parallel SparseMvMul, mgStepV with Chebyshev-Jacobi V(2,2) degree 4, solvePCG from 0 "
                     f"to 1e-8) on the same level systems; median of {len(runs)} solves",
           "cpu_model": CB.cpu_model(), "iterations": runs[-1][1],
           "true_rel_residual": None, "spmv_gbs": (12 * nnz + 4 * (n + 1) + 16 * n) / spmv_s / 1e9,
           "export_s": t_export, "setup_s": t_setup,
           "java": "not run: no JDK on the box (the reference is Java 11 with Maven dependencies)"}
    if x_gpu is not None:
        out["rel_l2_vs_gpu_fp64"] = float(np.linalg.norm(x - x_gpu) / np.linalg.norm(x_gpu))
    rp, ci, v = levels[-1]
    A = sps.csr_matrix((v, ci, rp), shape=(n, n))
    out["true_rel_residual"] = float(np.linalg.norm(rhs - A @ x) / np.linalg.norm(rhs))
    return out


def cpu_baseline(seconds_budget=5.0):
    """Oracle (single-threaded C restatement) MG-PCG on a bounded sample: 3D Q2
    16^3 cells (35,937 DOFs), same smoother and cycle; the solve is repeated for
    about seconds_budget of CPU time and the median is scaled by the DOF ratio.
    The sample's matrices are far smaller than C2's, so caches flatter the CPU.
    Reported beside the full-size baseline as a one-core figure."""
    from oracle import oracle as O
    from jfem_amd.workloads import sinusoid_terms
    base, nlev, k = [2, 2, 2], 4, 2
    cells = [16, 16, 16]
    t0 = time.perf_counter()
    G = O.MG(base, k, nlev, [0, 0, 0], [1, 1, 1], [O.GRAD_GRAD], [1.0], smoother=O.SM_CHEBYSHEV, runs=2,
             degree=4, lmin=[0.0] * nlev, lmax=[0.0] * nlev, cycles=1, fmg=False)
    for l in range(1, nlev):
        lmax = 1.1 * G.estimate_lmax(l)
        G.set_interval(l, lmax / 20.0, lmax)
    A = G.level_matrix(nlev - 1)
    kw, amp = sinusoid_terms(2, 16, 3)
    rhs = O.rhs_sines(cells, k, [0, 0, 0], [1, 1, 1], kw, amp)
    rhs[O.boundary_nodes(cells, k)] = 0
    setup = time.perf_counter() - t0
    solves = []
    while not solves or (sum(solves) < seconds_budget and len(solves) < 1000):
        t1 = time.perf_counter()
        x, it, res, _ = O.solve("cg", A, rhs, np.zeros(A.n), 1e-8 * np.linalg.norm(rhs), precond=G)
        solves.append(time.perf_counter() - t1)
    solve = float(np.median(solves))
    # SpMV throughput of the oracle on the sample matrix
    xx = np.random.default_rng(0).standard_normal(A.n)
    reps = 20
    t2 = time.perf_counter()
    for _ in range(reps):
        A.matvec(xx)
    t_spmv = (time.perf_counter() - t2) / reps
    bytes_spmv = 12 * A.nnz + 4 * (A.n + 1) + 16 * A.n
    return dict(solve_s=solve, repeats=len(solves), iterations=it, n=A.n, setup_s=setup,
                spmv_gbs=bytes_spmv / t_spmv / 1e9)


def measured_traffic(cells, name="pmc_spmv_fp64_q2_96.md"):
    """HBM bytes per launch from a committed rocprofv3 PMC summary (96^3 only)."""
    path = os.path.join(ROOT, "profiles", name)
    if cells != 96 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def copy_calibration(jf, n=1 << 27, reps=20):
    import ctypes as C
    from jfem_amd import _native as N
    ms = C.c_double()
    N.check(N.lib().jfem_copy_bench(jf.context().h, n, reps, C.byref(ms)))
    return 16 * n / (ms.value * 1e-3) / 1e9, ms.value


def read_calibration(jf, n=700 << 20, reps=20):
    """streaming read of n doubles (5.9 GB at the default, the size of the SpMV's streams)"""
    import ctypes as C
    from jfem_amd import _native as N
    ms = C.c_double()
    N.check(N.lib().jfem_read_bench(jf.context().h, n, reps, C.byref(ms)))
    return 8 * n / (ms.value * 1e-3) / 1e9, ms.value


def spmv_only(jf, cells):
    s = jf.ContinuousTPFESpace([0.0] * 3, [1.0] * 3, [cells] * 3)
    s.assembleCells()
    s.assembleFunctions(2)
    A = s.assemble_matrix([jf.TPCellIntegral("GradGrad")])
    n, _, nnz = A.info()
    x = jf.DenseVector(np.random.default_rng(0).standard_normal(n))
    y = jf.DenseVector(n)
    ms = A.spmv_time_ms(x, y, reps=20)
    gbs, cms = copy_calibration(jf)
    print(json.dumps({"spmv_fp64_ms": ms, "spmv_fp64_gbs": (12 * nnz + 20 * n + 4) / (ms * 1e-3) / 1e9,
                      "copy_gbs": gbs, "copy_ms": cms, "n": n, "nnz": nnz}), flush=True)


def assembly_only(jf, cells, reps=2):
    """jfem_assemble of the 3D Q2 GRAD_GRAD matrix on cells^3, `reps` times with a constant and
    `reps` times with a log-normal per-cell coefficient (host wall time per call, device-resident
    result), then the copy calibration kernel; for kernel traces and PMC runs"""
    s = jf.ContinuousTPFESpace([0.0] * 3, [1.0] * 3, [cells] * 3)
    s.assembleCells()
    s.assembleFunctions(2)
    coeff = np.exp(np.random.default_rng(0).standard_normal(cells ** 3))
    ctx = jf.context()
    out = {}
    for name, cc in (("constant", None), ("cell_coeff", coeff)):
        times = []
        for _ in range(reps):
            ctx.sync()
            t = time.perf_counter()
            A = s.assemble_matrix([jf.TPCellIntegral("GradGrad")], cell_coeff=cc)
            ctx.sync()
            times.append(time.perf_counter() - t)
            n, _, nnz = A.info()
            del A
        out[f"assembly_{name}_s"] = times
    gbs, cms = copy_calibration(jf, reps=3)
    print(json.dumps({**out, "n": n, "nnz": nnz, "copy_gbs": gbs}), flush=True)


def network_only(jf, cells, reps=5):
    """the learned smoother's network alone on the finest node lattice of `cells`^3 Q2 cells
    (2 cells + 1 nodes per axis), with the C2 network's shape (3D, 5 input channels) and random
    weights and input: `reps` forward passes through jfem_nn_forward_host, for kernel traces"""
    n = 2 * cells + 1
    net = jf.SmootherNet(3, 5, seed=3)
    rng = np.random.default_rng(0)
    net.set_params((rng.standard_normal(net.num_params()) * 0.05).astype(np.float32))
    x = rng.standard_normal((n * n * n, 5)).astype(np.float32)
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        y = net.forward([n, n, n], x.reshape(-1))
        times.append(time.perf_counter() - t)
    print(json.dumps({"network_forward_s_host_copies_included": times, "nodes": n ** 3,
                      "output_abs_sum": float(np.abs(y).sum())}), flush=True)


def c4_hierarchy(jf, field, precision, cells=1024, base=16, galerkin=False):
    nlev = 1
    c = cells
    while c > base:
        c //= 2
        nlev += 1
    spaces = []
    for l in range(nlev):
        sp = jf.ContinuousTPFESpace([0.0, 0.0], [1.0, 1.0], [base << l, base << l])
        sp.assembleCells()
        sp.assembleFunctions(1)
        spaces.append(sp)
    G = jf.MGPreconditioner(spaces, [jf.TPCellIntegral("GradGrad")], jf.ChebyshevSmoother(4, 1), cycles=1, fmg=False,
                            precision=precision, cell_coeff=np.exp(field), galerkin=galerkin)
    return G, spaces


def bench_c4(jf, args):
    """2D -div(a grad u) = f, Q1 on 1024^2 cells (1,050,625 DOFs), a = exp(g) with g a
    squared-exponential Gaussian random field (length 0.1, contrast 1e3).  The
    learned smoother (4 conv layers, 32 channels, BF16 MFMA; one application per
    smoothing call) is trained with the two-grid loss on (residual, error) pairs
    of V-cycles on training fields (seeds 4000+; every second right-hand side
    advances with the learned V-cycle itself), then
    evaluated on held-out fields (seeds 5000+): time to ||r|| <= 1e-8 ||b|| of
    (a) FP64 MG-PCG with the Chebyshev-Jacobi V-cycle and (b) FP64 flexible CG
    around the FP32 V-cycle with the learned smoother (Chebyshev fallback)."""
    from jfem_amd.workloads import gaussian_random_field, sinusoid_terms
    cells = 1024
    net = jf.SmootherNet(2, 2, seed=7)
    t0 = time.perf_counter()
    losses = []
    for f in range(args.train_fields):
        field = gaussian_random_field(4000 + f, (cells, cells), 0.1, 1e3)
        G, spaces = c4_hierarchy(jf, field, jf.FP64, galerkin=args.galerkin)
        G.set_learned(net, field, runs=1, fallback_ratio=0.5)
        h = G.train(rhs_per_field=max(1, args.pairs // 10), cycles_per_rhs=10, lr=3e-4, seed=f + 1, mode=5)
        losses.append(float(np.mean(h[-8:])))
        log(f"train field {f}: two-grid energy ratio {np.mean(h[:8]):.4f} -> {np.mean(h[-8:]):.4f}")
        del G
    t_train = time.perf_counter() - t0
    res = {"pcg64": [], "fcg32_learned": [], "fcg32_learned_both": [], "fcg32_cheb": [], "vcyc_cheb": [],
           "vcyc_learned": [], "err": [], "fallbacks": 0}
    kw, amp = sinusoid_terms(4, 8, 2)
    for f in range(args.eval_fields):
        field = gaussian_random_field(5000 + f, (cells, cells), 0.1, 1e3)
        G64, spaces = c4_hierarchy(jf, field, jf.FP64, galerkin=args.galerkin)
        fine = spaces[-1]
        fine.initializeRhs()
        fine.writeCellIntegralsToRhs([jf.TPRightHandSideIntegral(jf.SinusoidSource(kw, amp))], fine.getRhs())
        r = fine.getRhs().numpy()
        r[fine.boundary_nodes()] = 0
        b = jf.DenseVector(r)
        tol = 1e-8 * np.linalg.norm(r)
        A = G64.getFinestSystem()
        it = jf.IterativeSolver()
        x64 = it.solvePCG(A, G64, b, tol, x0=jf.DenseVector(len(r)))
        res["pcg64"].append((it.info.seconds, it.iterations))
        info, _ = G64.solve(b, jf.DenseVector(len(r)), tol, 200)
        res["vcyc_cheb"].append((info.seconds, info.iterations))
        G64.set_learned(net, field, runs=1, fallback_ratio=0.5)
        info, _ = G64.solve(b, jf.DenseVector(len(r)), tol, 200)
        res["vcyc_learned"].append((info.seconds, info.iterations))
        res["fallbacks"] += int(G64.learned_status()[1])
        del G64
        G32, _ = c4_hierarchy(jf, field, jf.FP32, galerkin=args.galerkin)
        G32.set_learned(net, field, runs=1, fallback_ratio=0.5)
        G32.set_learned_sides(args.learned_sides)
        it2 = jf.IterativeSolver()
        x32 = it2.solveFCG(G32.getFinestSystem(), G32, b, tol)
        res["fcg32_learned"].append((it2.info.seconds, it2.iterations))
        res["fallbacks"] += int(G32.learned_status()[1])
        if args.learned_sides != 3:
            # the network on both sides of the V-cycle, for comparison
            G32.set_learned_sides(3)
            G32.set_learned_active(True)
            it2b = jf.IterativeSolver()
            it2b.solveFCG(G32.getFinestSystem(), G32, b, tol)
            res["fcg32_learned_both"].append((it2b.info.seconds, it2b.iterations))
            res["fallbacks"] += int(G32.learned_status()[1])
        G32.set_learned_active(False)      # the same FP32 hierarchy with Chebyshev on every level
        it3 = jf.IterativeSolver()
        it3.solveFCG(G32.getFinestSystem(), G32, b, tol)
        res["fcg32_cheb"].append((it3.info.seconds, it3.iterations))
        res["err"].append(float(np.linalg.norm(x32.numpy() - x64.numpy()) / np.linalg.norm(x64.numpy())))
        log(f"eval field {f}: pcg64 {res['pcg64'][-1]}, fcg32+learned {res['fcg32_learned'][-1]}, "
            f"V cheb {res['vcyc_cheb'][-1]}, V learned {res['vcyc_learned'][-1]}")
        del G32
    med = lambda k, i: float(np.median([v[i] for v in res[k]]))
    value = med("fcg32_learned", 0)
    line = {
        "metric": "time-to-1e-8 residual (s): learned+FP32 MG-FCG vs FP64 MG-PCG, 2D Q1 1024^2 log-normal coefficient",
        "value": value, "unit": "s", "n_gpus": 1, "steps": args.eval_fields, "warmup": 0,
        "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64 outer / f32 multigrid / bf16 network", "data": "synthetic",
        "config": {"workload": "2D -div(a grad u)=f, Q1 1024^2 cells, a=exp(GRF) contrast 1e3"
                               + ("; Galerkin coarse systems" if args.galerkin else ""),
                   "train_fields": args.train_fields, "pairs_per_field": max(1, args.pairs // 10) * 10,
                   "eval_fields": args.eval_fields,
                   "learned_sides": {1: "pre-smoothing", 2: "post-smoothing", 3: "both"}[args.learned_sides]},
        "fp64_mgpcg_s": med("pcg64", 0), "fp64_mgpcg_iterations": med("pcg64", 1),
        "learned_fcg32_iterations": med("fcg32_learned", 1),
        **({"learned_both_sides_fcg32_s": med("fcg32_learned_both", 0),
            "learned_both_sides_fcg32_iterations": med("fcg32_learned_both", 1)} if res["fcg32_learned_both"] else {}),
        "fp32_cheb_fcg_s": med("fcg32_cheb", 0), "fp32_cheb_fcg_iterations": med("fcg32_cheb", 1),
        "learned_vs_best_plain_fp32": med("fcg32_cheb", 0) / value,
        "speedup_vs_fp64_mgpcg": med("pcg64", 0) / value,
        "vcycles_chebyshev": med("vcyc_cheb", 1), "vcycles_learned": med("vcyc_learned", 1),
        "vcycle_solve_s_chebyshev": med("vcyc_cheb", 0), "vcycle_solve_s_learned": med("vcyc_learned", 0),
        "fallbacks": res["fallbacks"], "max_rel_l2_vs_fp64": float(np.max(res["err"])),
        "train_s": t_train, "final_train_loss": losses[-1] if losses else None,
    }
    print(json.dumps(line), flush=True)


def c5_velocity_mg(jf, cells, nu, precision, smoother, base=16, ctx=None):
    """scalar Q2 hierarchy base^2 ... cells^2 on the unit square with GRAD_GRAD weight nu (a
    coarsest level of 16^2 cells measured fastest: 4^2 and 8^2 add launch-bound levels, 32^2
    makes the dense coarse inverse 4225^2)"""
    spaces = []
    c = base
    while c <= cells:
        sp = jf.ContinuousTPFESpace([0.0, 0.0], [1.0, 1.0], [c, c], ctx)
        sp.assembleCells()
        sp.assembleFunctions(2)
        spaces.append(sp)
        c *= 2
    G = jf.MGPreconditioner(spaces, [jf.TPCellIntegral(nu, "GradGrad")], smoother, cycles=1, fmg=False,
                            precision=precision)
    return G, spaces


def c5_system(jf, cells, nu):
    """Taylor-Hood Q2-Q1 on cells^2: lid-driven cavity (u = (1, 0) on y = 1 between the
    corners, 0 on the other walls) plus a seeded sinusoidal body force, first pressure pinned"""
    from jfem_amd.workloads import sinusoid_terms
    s = jf.TaylorHoodSpace([0.0, 0.0], [1.0, 1.0], [cells, cells])
    s.assembleCells()
    s.assembleFunctions(1)
    ints = [jf.MixedCellIntegral.fromVelocityIntegral(jf.TPVectorCellIntegral(nu, "GradGrad")),
            jf.MixedTPCellIntegral(-1.0, "DivValue")]
    jf.context().sync()
    t0 = time.perf_counter()
    A = s.writeCellIntegralsToMatrix(ints)
    jf.context().sync()
    t_asm = time.perf_counter() - t0
    comps = []
    for c in range(2):
        kw, amp = sinusoid_terms(60 + c, 8, 2)
        comps.append(jf.TrigSource(kw, amp))
    rhs = jf.DenseVector(s.ndofs())
    s.writeCellIntegralsToRhs([jf.MixedRightHandSideIntegral.fromVelocityIntegral(
        jf.TPVectorRightHandSideIntegral(jf.VectorTrigSource(comps), "Value"))], rhs)

    def lid(x):
        return (1.0 if abs(x[1] - 1.0) < 1e-12 and 0.0 < x[0] < 1.0 else 0.0, 0.0)

    s.writeVelocityBoundaryValuesTo(lid, A, rhs)
    s.overWriteValue(s.getVelocitySize(), 0.0, A, rhs)
    return s, A, rhs, t_asm


def bench_c5(jf, args):
    """2D Stokes, Taylor-Hood Q2-Q1 on 256^2 cells (526,338 velocity + 66,049 pressure
    unknowns), viscosity 1, lid-driven cavity plus a seeded sinusoidal body force, the
    first pressure pinned to 0.  FP64 flexible GMRES(50) to ||r|| <= 1e-8 ||b|| with the
    block lower-triangular preconditioner: velocity block = one FP32 V(1,1) cycle per
    component (scalar Q2 hierarchy 16^2 ... 256^2, Chebyshev-Jacobi degree 3; the second
    component's cycle on its own context and stream, concurrent with the first), with the
    learned smoother on the finest level (network trained on the 128^2 hierarchy) or with
    Chebyshev there too; Schur complement = pressure mass / viscosity, inverted by
    --mass-its Chebyshev steps preconditioned by the lumped mass (0: lumped inverse).  The velocity
    error is measured against an FP64 reference solve (FP64 multigrid, FGMRES to 1e-12)."""
    cells, nu, tol = 256, 1.0, 1e-8
    s, A, rhs, t_asm = c5_system(jf, cells, nu)
    n, _, nnz = A.info()
    log(f"C5 Taylor-Hood {cells}^2: n={n} nnz={nnz}, assembly {t_asm * 1e3:.1f} ms")
    cheb = jf.ChebyshevSmoother(degree=3, runs=1, ratio=20.0)
    nv2 = s.getVelocitySize()

    def solve(G, rtol, reps, mass_its=args.mass_its, comps=()):
        prec = jf.StokesBlockPreconditioner(s, G, nu, mass_its=mass_its, component_mgs=comps)
        out = []
        for _ in range(reps):
            x = jf.DenseVector(n)
            info, hist = jf.solveStokesFGMRES(A, prec, rhs, x, rtol, restart=50, maxit=2000)
            out.append((info.seconds, info.iterations, info.resnorm))
        return x, out

    G64, _ = c5_velocity_mg(jf, cells, nu, jf.FP64, cheb)
    xref, r = solve(G64, 1e-12, 1)
    log(f"FP64 reference solve: {r[0][1]} its, {r[0][0]:.3f}s")
    xref_u = xref.numpy()[:nv2]
    fp64 = solve(G64, tol, args.warmup + args.steps)[1][args.warmup:]
    log(f"FP64 multigrid FGMRES to 1e-8: {fp64[-1][1]} its, {np.median([v[0] for v in fp64]):.3f}s")
    del G64
    G32, _ = c5_velocity_mg(jf, cells, nu, jf.FP32, cheb)
    # the second velocity component's V-cycle runs on its own context (stream), concurrently
    ctx2 = jf.Context(0)
    G32b, _ = c5_velocity_mg(jf, cells, nu, jf.FP32, cheb, ctx=ctx2)
    seq = solve(G32, tol, args.warmup + args.steps)[1][args.warmup:]
    log(f"FP32 Chebyshev multigrid FGMRES, components in sequence: {seq[-1][1]} its, "
        f"{np.median([v[0] for v in seq]):.3f}s")
    x32, fp32 = solve(G32, tol, args.warmup + args.steps, comps=(G32b,))
    fp32 = fp32[args.warmup:]
    err32 = float(np.linalg.norm(x32.numpy()[:nv2] - xref_u) / np.linalg.norm(xref_u))
    log(f"FP32 Chebyshev multigrid FGMRES: {fp32[-1][1]} its, {np.median([v[0] for v in fp32]):.3f}s")
    lumped = solve(G32, tol, args.warmup + args.steps, 0, comps=(G32b,))[1][args.warmup:]
    log(f"... with the lumped-mass Schur complement: {lumped[-1][1]} its, {np.median([v[0] for v in lumped]):.3f}s")
    line_extra = {}
    value = float(np.median([v[0] for v in fp32]))
    if not args.no_learned:
        t0 = time.perf_counter()
        Gt, _ = c5_velocity_mg(jf, 128, nu, jf.FP64, cheb)
        net = jf.SmootherNet(2, args.c5_cin, seed=7)
        Gt.set_learned(net, np.zeros(128 * 128), runs=1, fallback_ratio=0.5)
        h = Gt.train(rhs_per_field=400, cycles_per_rhs=8, lr=1e-4, seed=1, mode=5)
        jf.context().sync()
        t_train = time.perf_counter() - t0
        del Gt
        log(f"learned smoother trained on 128^2 in {t_train:.1f}s, two-grid energy ratio "
            f"{np.mean(h[:8]):.3f} -> {np.mean(h[-8:]):.4f}")
        G32.set_learned(net, np.zeros(cells * cells), runs=1, fallback_ratio=0.5)
        net2 = jf.SmootherNet(2, args.c5_cin, ctx=ctx2)   # the same weights for the second component's hierarchy
        net2.set_params(net.get_params())
        G32b.set_learned(net2, np.zeros(cells * cells), runs=1, fallback_ratio=0.5)

        def solve_learned(sides):
            for G_ in (G32, G32b):
                G_.set_learned_sides(sides)
            out = []
            for _ in range(args.warmup + args.steps):
                G32.set_learned_active(True)
                G32b.set_learned_active(True)
                x_, r_ = solve(G32, tol, 1, comps=(G32b,))
                out.append(r_[0])
            return x_, out[args.warmup:]

        both = solve_learned(3)[1] if args.learned_sides != 3 else None
        xl, res = solve_learned(args.learned_sides)
        errl = float(np.linalg.norm(xl.numpy()[:nv2] - xref_u) / np.linalg.norm(xref_u))
        value = float(np.median([v[0] for v in res]))
        fell = bool(G32.learned_status()[1] or G32b.learned_status()[1])
        log(f"learned + FP32 multigrid FGMRES: {res[-1][1]} its, {value:.3f}s, fell back {fell}")
        line_extra = {"learned_sides": {1: "pre-smoothing", 2: "post-smoothing", 3: "both"}[args.learned_sides],
                      "learned_fgmres_iterations": res[-1][1], "learned_velocity_rel_l2_vs_fp64": errl,
                      "learned_fell_back": fell, "train_s": t_train,
                      "final_train_loss": float(np.mean(h[-8:]))}
        if both:
            line_extra["learned_both_sides_s"] = float(np.median([v[0] for v in both]))
            line_extra["learned_both_sides_fgmres_iterations"] = both[-1][1]
    med = lambda r, i: float(np.median([v[i] for v in r]))
    line = {
        "metric": "time-to-1e-8 residual (s): Taylor-Hood Q2-Q1 Stokes FGMRES(50), block preconditioner with "
                  "learned + FP32 velocity multigrid",
        "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64 outer / f32 multigrid / bf16 network", "data": "synthetic",
        "config": {"workload": "2D Stokes, Taylor-Hood Q2-Q1 256^2 cells, lid-driven + sinusoidal body force, nu=1",
                   "unknowns": n, "nnz": nnz},
        "th_assembly_s": t_asm,
        "fp64_mg_fgmres_s": med(fp64, 0), "fp64_mg_fgmres_iterations": fp64[-1][1],
        "fp32_cheb_mg_fgmres_s": med(fp32, 0), "fp32_cheb_mg_fgmres_iterations": fp32[-1][1],
        "fp32_cheb_velocity_rel_l2_vs_fp64": err32,
        "fp32_cheb_components_in_sequence_s": med(seq, 0),
        "schur_mass_chebyshev_steps": args.mass_its,
        "fp32_cheb_lumped_schur_s": med(lumped, 0), "fp32_cheb_lumped_schur_iterations": lumped[-1][1],
    }
    line.update(line_extra)
    print(json.dumps(line), flush=True)


def bench_c3(jf, args):
    """3D -div(a grad u) = f, Q2 on 96^3 cells, a = exp(g) per cell with g a seeded
    squared-exponential Gaussian random field (length 0.1, contrast 1e4); the
    coarser levels average the coefficient of their 2^3 children.  Time to
    ||r|| <= 1e-8 ||b|| of FP64 MG-PCG and of the FP32 multigrid inside FP64
    flexible CG (the FP64 residual decides convergence), and the relative L2
    difference of the two solutions."""
    from jfem_amd.workloads import gaussian_random_field
    nlev, c = 1, args.cells
    while c % 2 == 0 and c > 3:
        c //= 2
        nlev += 1
    base = [c, c, c]
    g = gaussian_random_field(3, (args.cells,) * 3, 0.1, 1e4)
    coeff = np.exp(g)
    smoother = jf.ChebyshevSmoother(degree=4, runs=2, ratio=20.0)
    spaces, G, t_mg = build_problem(jf, base, nlev, 2, smoother, jf.FP64, 3, 16, cell_coeff=coeff)
    fine = spaces[-1]
    A = G.getFinestSystem()
    n, _, nnz = A.info()
    b, rhs = rhs_for(jf, fine, 3, 16)
    bnorm = float(np.linalg.norm(rhs))
    tol = 1e-8 * bnorm
    ctx = jf.context()

    def timed(solve):
        for _ in range(args.warmup):
            solve()
        ctx.sync()
        t0 = time.perf_counter()
        ts = []
        for _ in range(args.steps):
            it, x = solve()
            ts.append(it.info.seconds)
        ctx.sync()
        return it, x, float(np.median(ts)), (time.perf_counter() - t0) / args.steps * 1e3

    def pcg64():
        it = jf.IterativeSolver()
        x = it.solvePCG(A, G, b, tol, x0=jf.DenseVector(n))
        return it, x

    it64, x64, t64, ms64 = timed(pcg64)
    log(f"FP64 MG-PCG: {it64.iterations} its, {t64:.3f}s")
    x64 = x64.numpy()
    del G
    _, G32, _ = build_problem(jf, base, nlev, 2, smoother, jf.FP32, 3, 16, cell_coeff=coeff)
    A32 = G32.getFinestSystem()

    def fcg32():
        it = jf.IterativeSolver()
        x = it.solveFCG(A32, G32, b, tol)
        return it, x

    it32, x32, t32, ms32 = timed(fcg32)
    log(f"FP32 MG in FP64 FCG: {it32.iterations} its, {t32:.3f}s")
    del G32
    # Galerkin coarse systems (P^T A P from the finest one, AMGFromMatrix) instead of
    # re-assembling the averaged coefficient on every level
    t0 = time.perf_counter()
    _, Gg, _ = build_problem(jf, base, nlev, 2, smoother, jf.FP64, 3, 16, cell_coeff=coeff, galerkin=True)
    ctx.sync()
    t_gal = time.perf_counter() - t0

    def pcg64g():
        it = jf.IterativeSolver()
        x = it.solvePCG(A, Gg, b, tol, x0=jf.DenseVector(n))
        return it, x

    it64g, _, t64g, _ = timed(pcg64g)
    log(f"FP64 MG-PCG, Galerkin coarse systems: {it64g.iterations} its, {t64g:.3f}s")
    del Gg
    _, Gg32, _ = build_problem(jf, base, nlev, 2, smoother, jf.FP32, 3, 16, cell_coeff=coeff, galerkin=True)
    A32g = Gg32.getFinestSystem()

    def fcg32g():
        it = jf.IterativeSolver()
        x = it.solveFCG(A32g, Gg32, b, tol)
        return it, x

    it32g, x32g, t32g, ms32g = timed(fcg32g)
    log(f"FP32 MG (Galerkin) in FP64 FCG: {it32g.iterations} its, {t32g:.3f}s")
    r = A.mvMul(x32g)
    line = {
        "metric": "time-to-1e-8 residual (s): FP32 MG-FCG vs FP64 MG-PCG, 3D Q2 heterogeneous coefficient",
        "value": t32g, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms32g,
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f64 outer / f32 multigrid",
        "data": "synthetic",
        "config": {"workload": f"3D -div(a grad u)=f, Q2 {args.cells}^3 cells, a=exp(GRF) length 0.1 contrast 1e4; "
                               f"FP32 GMG with Galerkin coarse systems inside FP64 flexible CG vs FP64 GMG "
                               f"(re-assembled coarse systems) PCG",
                   "ndofs": n, "nnz": nnz, "levels": nlev},
        "fp64_mgpcg_s": t64, "fp64_mgpcg_iterations": it64.iterations, "fp64_ms_per_step": ms64,
        "fp32mg_fcg_s": t32, "fp32mg_fcg_iterations": it32.iterations,
        "fp64_mgpcg_galerkin_s": t64g, "fp64_mgpcg_galerkin_iterations": it64g.iterations,
        "fp32mg_galerkin_fcg_iterations": it32g.iterations,
        "speedup_fp32mg_galerkin_vs_fp64": t64 / t32g, "speedup_fp32mg_vs_fp64": t64 / t32,
        "fp32mg_galerkin_true_rel_residual": float(np.linalg.norm(rhs - r.numpy()) / bnorm),
        "fp32mg_galerkin_vs_fp64_rel_l2": float(np.linalg.norm(x32g.numpy() - x64) / np.linalg.norm(x64)),
        "fp32mg_vs_fp64_rel_l2": float(np.linalg.norm(x32.numpy() - x64) / np.linalg.norm(x64)),
        "hierarchy_setup_s": t_mg, "galerkin_hierarchy_setup_s": t_gal,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=96, help="cells per axis of the finest grid (3 * 2^m)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-smoother solve")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="c2: 3D Q2 96^3 learned + FP32 MG vs FP64 MG-PCG (default); c3: the same grid with a "
                         "log-normal coefficient (contrast 1e4), FP32 MG in FP64 FCG vs FP64 MG-PCG; c4: 2D Q1 "
                         "1024^2 log-normal coefficient, learned smoother vs Chebyshev; c5: 2D Taylor-Hood Stokes "
                         "256^2, block-preconditioned FGMRES")
    ap.add_argument("--train-fields", type=int, default=200, help="C4: training coefficient fields (seeds 4000+)")
    ap.add_argument("--pairs", type=int, default=100, help="C4: (residual, error) pairs per training field")
    ap.add_argument("--eval-fields", type=int, default=50, help="C4: held-out evaluation fields (seeds 5000+)")
    ap.add_argument("--galerkin", action="store_true", help="C4: Galerkin coarse systems on every hierarchy")
    ap.add_argument("--c5-cin", type=int, default=4, choices=[2, 4],
                    help="C5: network input channels (2: residual and coefficient; 4: plus node-parity channels)")
    ap.add_argument("--mass-its", type=int, default=3,
                    help="C5: Chebyshev steps on the pressure mass matrix in the Schur complement (0 = lumped)")
    ap.add_argument("--learned-sides", type=int, default=None, choices=[1, 2, 3],
                    help="V-cycle smoothing calls on the finest level that use the network (1 pre, 2 post, 3 both); "
                         "the other side keeps Chebyshev.  Default: the fastest measured per workload (C2 and C4: "
                         "2, C5: 3)")
    ap.add_argument("--network-only", action="store_true",
                    help="time the learned smoother's network forward alone on the finest node lattice")
    ap.add_argument("--assembly-only", action="store_true",
                    help="assemble the finest 3D Q2 matrix with a constant and a per-cell coefficient (for kernel "
                         "traces and PMC runs)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="assemble the finest matrix and time only the SpMV and the copy calibration (for PMC runs)")
    args = ap.parse_args()
    if args.learned_sides is None:
        args.learned_sides = 3 if args.workload == "c5" else 2
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solve path runs on a single GPU "
              "(a multi-GPU slab decomposition is designed but not implemented); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import jfem_amd as jf

    if args.network_only:
        network_only(jf, args.cells)
        return
    if args.spmv_only:
        spmv_only(jf, args.cells)
        return
    if args.assembly_only:
        assembly_only(jf, args.cells)
        return
    if args.workload == "c4":
        bench_c4(jf, args)
        return
    if args.workload == "c3":
        bench_c3(jf, args)
        return
    if args.workload == "c5":
        bench_c5(jf, args)
        return

    nlev = 1
    c = args.cells
    while c % 2 == 0 and c > 3:
        c //= 2
        nlev += 1
    base = [c, c, c]
    log(f"building hierarchy {base} x 2^{nlev - 1}")
    smoother = jf.ChebyshevSmoother(degree=4, runs=2, ratio=20.0)
    spaces, G, t_mg = build_problem(jf, base, nlev, 2, smoother, jf.FP64, 2, 16)
    fine = spaces[-1]
    A = G.getFinestSystem()
    n, _, nnz = A.info()
    log(f"n={n} nnz={nnz} hierarchy setup {t_mg:.2f}s")

    # assembly of the finest matrix alone (jfem_assemble), timed
    ctx = jf.context()
    ctx.sync()
    t0 = time.perf_counter()
    A_asm = fine.assemble_matrix([jf.TPCellIntegral("GradGrad")])
    ctx.sync()
    t_asm = time.perf_counter() - t0
    del A_asm
    log(f"finest assembly {t_asm * 1e3:.1f} ms")

    b, rhs = rhs_for(jf, fine, 2, 16)
    bnorm = float(np.linalg.norm(rhs))
    tol = 1e-8 * bnorm

    def solve_fp64():
        x = jf.DenseVector(n)
        it = jf.IterativeSolver()
        it.solvePCG(A, G, b, tol, x0=x)
        return it, x

    for w in range(args.warmup):
        it, x = solve_fp64()
        log(f"warmup {w}: {it.iterations} its, {it.info.seconds:.3f}s")
    times, iters = [], []
    ctx.sync()
    t_start = time.perf_counter()
    for s in range(args.steps):
        it, x = solve_fp64()
        times.append(it.info.seconds)
        iters.append(it.iterations)
        log(f"step {s}: {it.iterations} its, {it.info.seconds:.3f}s, resnorm/b {it.info.resnorm / bnorm:.2e}")
    ctx.sync()
    wall = time.perf_counter() - t_start
    ms_per_step = wall / args.steps * 1e3
    # true residual of the last solution
    r = A.mvMul(x)
    true_rel = float(np.linalg.norm(rhs - r.numpy()) / bnorm)
    x64 = x.numpy()

    cpu = None
    if not args.no_cpu:
        log("CPU baseline: OpenMP restatement at full size on all host cores")
        cpu = cpu_baseline_full(G, nlev, base, 2, rhs, tol, x_gpu=x64)
        log("CPU oracle, one core, bounded sample")
        cb = cpu_baseline()
        scale = n / cb["n"]
        cpu["oracle_1core_sample"] = (f"oracle MG-PCG on 3D Q2 16^3 cells ({cb['n']} DOFs), {cb['iterations']} its in "
                                      f"{cb['solve_s']:.3f}s (median of {cb['repeats']} solves)")
        cpu["oracle_1core_extrapolated_s"] = cb["solve_s"] * scale

    # roofline of the FP64 SpMV on the finest level
    xs = jf.DenseVector(np.random.default_rng(0).standard_normal(n))
    ys = jf.DenseVector(n)
    # steady state: one untimed batch (the GPU idled during the CPU baseline), then the median of
    # three timed batches of 50 launches each
    A.spmv_time_ms(xs, ys, reps=20)
    spmv_ms = float(np.median([A.spmv_time_ms(xs, ys, reps=50) for _ in range(3)]))
    spmv_bytes = 12 * nnz + 4 * (n + 1) + 16 * n
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    A.spmv_time_ms(xs, ys, reps=20, fp32=True)
    spmv32_ms = float(np.median([A.spmv_time_ms(xs, ys, reps=50, fp32=True) for _ in range(3)]))
    spmv32_gbs = (8 * nnz + 4 * (n + 1) + 8 * n) / (spmv32_ms * 1e-3) / 1e9
    log(f"SpMV fp64 {spmv_ms:.3f} ms {spmv_gbs:.0f} GB/s; fp32 {spmv32_ms:.3f} ms {spmv32_gbs:.0f} GB/s")
    # the solve's dominant launch: the finest level's Chebyshev step (matrix-free), timed the same way;
    # algorithmic bytes: x, b, D^-1 and d read, the new x and d written, 8 bytes each per row (the
    # constant coefficient adds none)
    G.level_step_time_ms(nlev - 1, xs, ys, reps=20)
    step_ms = float(np.median([G.level_step_time_ms(nlev - 1, xs, ys, reps=50) for _ in range(3)]))
    step_bytes = 48 * n
    step_gbs = step_bytes / (step_ms * 1e-3) / 1e9
    log(f"finest Chebyshev step {step_ms * 1e3:.1f} us {step_gbs:.0f} GB/s")
    copy_gbs, _ = copy_calibration(jf)
    read_gbs, _ = read_calibration(jf)
    log(f"copy calibration {copy_gbs:.0f} GB/s, streaming read {read_gbs:.0f} GB/s")

    extra = {}
    if not args.no_fp32:
        del G
        _, G32, t_mg32 = build_problem(jf, base, nlev, 2, smoother, jf.FP32, 2, 16)
        A32 = G32.getFinestSystem()
        t32 = []
        for s in range(max(1, args.steps)):
            it = jf.IterativeSolver()
            x32 = it.solveFCG(A32, G32, b, tol)
            t32.append(it.info.seconds)
            log(f"fp32-MG FCG: {it.iterations} its, {it.info.seconds:.3f}s")
        err = float(np.linalg.norm(x32.numpy() - x64) / np.linalg.norm(x64))
        extra = {"fp32mg_fcg_s": float(np.median(t32)), "fp32mg_fcg_iterations": it.iterations,
                 "fp32mg_vs_fp64_rel_l2": err, "speedup_fp32mg_vs_fp64": float(np.median(times) / np.median(t32))}
        if not args.no_learned:
            # learned smoother: trained on the same Q2 Poisson operator on 24^3 cells (the
            # network input is the Jacobi-scaled residual, independent of the mesh size),
            # then attached to the finest level of the FP32 hierarchy (one application per
            # smoothing call, Chebyshev fallback when a cycle reduces less than 0.5x).  The
            # coarser levels smooth with Chebyshev-Jacobi deg 3 V(1,1), in training too.
            # 400 right-hand sides x 8 cycles at learning rate 1e-4: 5 FCG iterations for every
            # training seed tried (1, 2, 3); at 3e-4 the result depends on the seed (5 or 6),
            # 40 x 8 leaves a weaker smoother (7), and 1200 x 8 can diverge late
            coarse_sm = jf.ChebyshevSmoother(degree=3, runs=1, ratio=20.0)
            t0 = time.perf_counter()
            _, Gt, _ = build_problem(jf, base, 4, 2, coarse_sm, jf.FP64, 2, 16)
            net = jf.SmootherNet(3, 5, seed=7)
            # constant coefficient: no log-coefficient channel (it would be zero everywhere)
            Gt.set_learned(net, None, runs=1, fallback_ratio=0.5)
            h = Gt.train(rhs_per_field=400, cycles_per_rhs=8, lr=1e-4, seed=1, mode=5)
            ctx.sync()
            t_train = time.perf_counter() - t0
            del Gt
            log(f"learned smoother trained on 24^3 in {t_train:.1f}s, two-grid energy ratio "
                f"{np.mean(h[:8]):.3f} -> {np.mean(h[-8:]):.4f}")
            tl, fb = [], False

            # the learned hierarchy: the network smooths the finest level; the coarser levels keep
            # Chebyshev-Jacobi, degree 3, one application per side (V(1,1)), which measured fastest
            # for this pairing; the same hierarchy without the network is timed below
            del G32
            _, G32, _ = build_problem(jf, base, nlev, 2, coarse_sm, jf.FP32, 2, 16)
            A32 = G32.getFinestSystem()
            G32.set_learned(net, None, runs=1, fallback_ratio=0.5)
            # the network on one side of the V-cycle (1 = pre-smoothing, 2 = post-smoothing) or both
            # (3); the other side keeps the hierarchy's Chebyshev smoother
            for sides in (1, 2, 3):
                if sides == args.learned_sides:
                    continue
                G32.set_learned_sides(sides)
                ts_ = []
                for s in range(2):
                    G32.set_learned_active(True)
                    its_ = jf.IterativeSolver()
                    its_.solveFCG(A32, G32, b, tol)
                    ts_.append(its_.info.seconds)
                name = {1: "pre_only", 2: "post_only", 3: "both_sides"}[sides]
                extra[f"learned_{name}_fcg_s"] = float(np.median(ts_))
                extra[f"learned_{name}_fcg_iterations"] = its_.iterations
                log(f"learned smoother, {name.replace('_', ' ')}: {its_.iterations} its, {np.median(ts_):.3f}s")
            G32.set_learned_sides(args.learned_sides)

            def solve_learned():
                G32.set_learned_active(True)   # re-arms the smoother and clears the fallback flag
                itl = jf.IterativeSolver()
                xs_ = itl.solveFCG(A32, G32, b, tol)
                return itl, xs_, G32.learned_status()[1]

            for w in range(args.warmup):
                solve_learned()
            ctx.sync()
            t_start = time.perf_counter()
            for s in range(args.steps):
                it, xl, f_ = solve_learned()
                tl.append(it.info.seconds)
                fb = fb or f_
                log(f"learned + fp32-MG FCG: {it.iterations} its, {it.info.seconds:.3f}s, fell back {f_}")
            ctx.sync()
            extra["learned_ms_per_step"] = (time.perf_counter() - t_start) / args.steps * 1e3
            rl = A.mvMul(xl)
            # the same FP32 hierarchy with Chebyshev on every level (no network)
            G32.set_learned_active(False)
            tc = []
            for s in range(max(1, args.steps)):
                itc = jf.IterativeSolver()
                itc.solveFCG(A32, G32, b, tol)
                tc.append(itc.info.seconds)
            extra["fp32mg_cheb_v11_deg3_fcg_s"] = float(np.median(tc))
            extra["fp32mg_cheb_v11_deg3_fcg_iterations"] = itc.iterations
            log(f"same FP32 hierarchy, Chebyshev deg 3 V(1,1) on every level: {itc.iterations} its, "
                f"{np.median(tc):.3f}s")
            del G32
            # the fastest plain FP32 cycle found: degree-1 Chebyshev (damped Jacobi) V(1,1)
            _, G1, _ = build_problem(jf, base, nlev, 2, jf.ChebyshevSmoother(degree=1, runs=1, ratio=20.0),
                                     jf.FP32, 2, 16)
            A1 = G1.getFinestSystem()
            tc = []
            for s in range(max(1, args.steps) + 1):
                itc = jf.IterativeSolver()
                itc.solveFCG(A1, G1, b, tol)
                tc.append(itc.info.seconds)
            extra["fp32mg_cheb_v11_deg1_fcg_s"] = float(np.median(tc[1:]))
            extra["fp32mg_cheb_v11_deg1_fcg_iterations"] = itc.iterations
            log(f"FP32 hierarchy, Chebyshev deg 1 V(1,1): {itc.iterations} its, {np.median(tc[1:]):.3f}s")
            del G1
            # stationary V-cycle iteration with each smoother (FP64 hierarchy, same network)
            G64 = build_problem(jf, base, nlev, 2, smoother, jf.FP64, 2, 16)[1]
            G64.set_learned(net, None, runs=1, fallback_ratio=0.5)
            for name, active in (("chebyshev", False), ("learned", True)):
                G64.set_learned_active(active)
                G64.solve(b, jf.DenseVector(n), tol, 100)
                info, _ = G64.solve(b, jf.DenseVector(n), tol, 100)
                extra[f"vcycle_{name}_cycles"] = info.iterations
                extra[f"vcycle_{name}_s"] = info.seconds
                log(f"V-cycle iteration, {name}: {info.iterations} cycles, {info.seconds:.3f}s")
            del G64
            extra.update({
                "learned_fp32mg_fcg_s": float(np.median(tl)), "learned_fp32mg_fcg_iterations": it.iterations,
                "learned_fell_back": bool(fb), "learned_train_s": t_train,
                "learned_true_rel_residual": float(np.linalg.norm(rhs - rl.numpy()) / bnorm),
                "learned_vs_fp64_rel_l2": float(np.linalg.norm(xl.numpy() - x64) / np.linalg.norm(x64)),
                "speedup_learned_fp32mg_vs_fp64": float(np.median(times) / np.median(tl))})

    fp64_s = float(np.median(times))
    learned = "learned_fp32mg_fcg_s" in extra
    value = extra["learned_fp32mg_fcg_s"] if learned else fp64_s
    sides_txt = {1: "pre-smoothing side", 2: "post-smoothing side", 3: "both sides"}[args.learned_sides]
    workload = (f"3D Poisson Q2 {args.cells}^3 cells; learned smoother (finest level, {sides_txt}; Chebyshev-Jacobi "
                f"deg 3 on the other side and V(1,1) below) + FP32 GMG V-cycle inside FP64 flexible CG vs FP64 GMG "
                f"V(2,2) Chebyshev-Jacobi deg 4 PCG, both to 1e-8" if learned else
                f"3D Poisson Q2 {args.cells}^3 cells, FP64 GMG V(2,2) Chebyshev-Jacobi deg 4, CG to 1e-8")
    line = {
        "metric": "time-to-1e-8 residual (s): learned+FP32 MG-PCG vs FP64 MG-PCG; SpMV GB/s" if learned else
                  "time-to-1e-8 residual (s): FP64 MG-PCG",
        "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": extra.pop("learned_ms_per_step") if learned else ms_per_step, "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None,
        "dtype": "f64 outer / f32 multigrid / bf16 network" if learned else "f64", "data": "synthetic",
        "config": {"workload": workload, "ndofs": n, "nnz": nnz, "levels": nlev,
                   "fp64_iterations": int(np.median(iters))},
        "fp64_mgpcg_s": fp64_s, "fp64_ms_per_step": ms_per_step,
        "true_rel_residual": true_rel,
        "assembly_s": t_asm, "hierarchy_setup_s": t_mg,
        "spmv_fp64_ms": spmv_ms, "spmv_fp64_gbs": spmv_gbs, "spmv_fp32_ms": spmv32_ms, "spmv_fp32_gbs": spmv32_gbs,
        "roofline": {"bound": "hbm", "achieved": spmv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": spmv_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(args.cells),
                     "kernel": "k_stream<4096,double,double,EpiStore<double>> (finest-level CSR SpMV)",
                     "traffic_source": "profiles/pmc_spmv_fp64_q2_96.md: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 "
                                       "per launch, the factors from the copy kernel of known size",
                     "achieved_timing": "hipEvents on the context stream, median of 3 batches of 50 launches after "
                                        "a warm-up batch",
                     "bytes_per_launch": spmv_bytes, "copy_kernel_gbs": copy_gbs,
                     "read_kernel_gbs": read_gbs, "frac_of_read_kernel": spmv_gbs / read_gbs},
        "roofline_smoother": {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": step_gbs / HBM_PEAK_GBS,
                              "traffic": measured_traffic(args.cells, "pmc_matrix_free_q2_96.md"),
                              "kernel": "k_mf<double,3,2,EpiCheby<double>> (finest-level matrix-free Chebyshev "
                                        "step, the FP64 solve's dominant launch)",
                              "traffic_source": "profiles/pmc_matrix_free_q2_96.md: rocprofv3 FETCH_SIZE x 2.0 + "
                                                "WRITE_SIZE x 1.0 per launch",
                              "achieved_timing": "hipEvents on the context stream, median of 3 batches of 50 "
                                                 "launches after a warm-up batch",
                              "bytes_per_launch": step_bytes, "ms": step_ms},
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
