/*
 * jfem_hip.h -- C ABI of libjfem_hip.so, the MI355X (gfx950) implementation of
 * JFEM's solve path for tensor-product Qk spaces: structured space and DOF
 * numbering, TPCellIntegral assembly into CSR, Dirichlet elimination, the
 * sparse matrix-vector product, CG / BiCGStab / GMRES and the geometric
 * multigrid preconditioner.
 *
 * Conventions (mirroring the reference's Java API and its JNI sparse solver,
 * JSparse/src/com/simonschmidt/SparseSolver.java:22-46):
 *  - every entry point returns a jfem_status (0 = success); failures never
 *    throw across the ABI and jfem_last_error() returns the message;
 *  - host arrays are caller-owned and only read/written during the call;
 *  - jfem_vec / jfem_mat / jfem_mg handles own device (HBM) memory, so solves
 *    run with operands resident on the GPU;
 *  - one context per host thread; calls on one context are not re-entrant
 *    (as the reference's IterativeSolver instances are not).
 * Reference paths are relative to the JFEM repository, M/ =
 * Core/src/main/java/.
 */
#ifndef JFEM_HIP_H
#define JFEM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    JFEM_OK = 0,
    JFEM_ERR_ARG = 1,          /* IllegalArgumentException in the reference */
    JFEM_ERR_STATE = 2,        /* IllegalStateException */
    JFEM_ERR_UNSUPPORTED = 3,  /* UnsupportedOperationException */
    JFEM_ERR_DEVICE = 4,       /* HIP runtime failure */
    JFEM_ERR_NOMEM = 5
} jfem_status;

/* integral kinds: TPCellIntegral names (M/tensorproduct/TPCellIntegral.java:20-24) */
enum { JFEM_GRAD_GRAD = 0, JFEM_VALUE_VALUE = 1, JFEM_H1 = 2 };

/* quadrature rules: QuadratureRule1D (M/tensorproduct/QuadratureRule1D.java:11-64);
 * 1..5 = Gauss1..Gauss5 (Gauss5 is TPCellIntegral's default), 12..16 = GaussLobatto2..6 */
enum { JFEM_GAUSS5 = 5 };

/* smoothers */
enum { JFEM_SM_RICHARDSON = 0, JFEM_SM_JACOBI = 1, JFEM_SM_CHEBYSHEV = 2 };

/* precision of the multigrid preconditioner */
enum { JFEM_FP64 = 0, JFEM_FP32 = 1 };

typedef struct jfem_ctx jfem_ctx;
typedef struct jfem_space jfem_space;
typedef struct jfem_vec jfem_vec;
typedef struct jfem_mat jfem_mat;
typedef struct jfem_mg jfem_mg;
typedef struct jfem_nn jfem_nn;

const char *jfem_last_error(void);

/* ---- context ------------------------------------------------------------ */
/* no reference counterpart: owns the HIP stream (plus an auxiliary stream that assembly uses
 * internally and joins before it returns), reduction scratch and set-up scratch (kept until
 * jfem_ctx_destroy) */
int jfem_ctx_create(int device, jfem_ctx **out);
int jfem_ctx_destroy(jfem_ctx *ctx);
int jfem_ctx_sync(jfem_ctx *ctx);
/* event timer on the context's stream: start, then stop returns milliseconds */
int jfem_timer_start(jfem_ctx *ctx);
int jfem_timer_stop(jfem_ctx *ctx, double *ms);
/* HBM calibration (no reference counterpart): mean time of a plain copy of n
 * doubles (8n bytes read + 8n written) over `reps` launches */
int jfem_copy_bench(jfem_ctx *ctx, long n, int reps, double *ms);
/* calibration (not in the reference): average time of a streaming read of n
 * doubles (16-byte loads, 16 KiB per workgroup), for the achievable read
 * bandwidth that the SpMV roofline is compared with */
int jfem_read_bench(jfem_ctx *ctx, long n, int reps, double *ms);

/* ---- space: new ContinuousTPFESpace(start, end, cells) + assembleFunctions(k)
 *      (M/tensorproduct/ContinuousTPFESpace.java:17-50) -------------------- */
int jfem_space_create(jfem_ctx *ctx, int dim, const int *cells, const double *start, const double *end,
                      int degree, jfem_space **out);
int jfem_space_destroy(jfem_space *s);
long jfem_space_ndofs(const jfem_space *s);
/* global index of every (cell, local function): cells in IntCoordinates order,
 * local index with axis 0 fastest; out has ncells*(k+1)^dim ints */
int jfem_space_cell_dofs(jfem_space *s, int *out);
/* Dirichlet nodes of the whole boundary, ascending (getBoundaryNodes,
 * M/basic/AcceptsMatrixBoundaryValues.java:46-60); out has ndofs capacity */
int jfem_space_boundary_nodes(jfem_space *s, int *out, long *count);
/* node point of every DOF (LagrangeNodeFunctional point), out[dof*dim + a] */
int jfem_space_node_coords(jfem_space *s, double *out);
/* ScalarFESpaceFunction(functions, coeffs).value(p) (M/basic/ScalarFESpaceFunction.java:40-45)
 * at npts points (points[i*dim + a], host): sum of coefficient * shape-function value,
 * each shape function evaluated in the first of its cells containing the point
 * (ScalarShapeFunction.value, M/basic/ScalarShapeFunction.java:26-31); 0 outside
 * the domain.  values: npts doubles (host). */
int jfem_space_evaluate(jfem_space *s, const jfem_vec *coeffs, long npts, const double *points, double *values);
/* ConvergenceOrderEstimator.normL2Difference (Core/src/test/java/examples/
 * ConvergenceOrderEstimator.java:49-57) of two FE functions (spaces of the same
 * dimension, e.g. two refinement levels) at npts host points:
 * sqrt(sum_i (f1(p_i) - f2(p_i))^2). */
int jfem_fe_l2_difference(jfem_space *s1, const jfem_vec *c1, jfem_space *s2, const jfem_vec *c2, long npts,
                          const double *points, double *result);

/* ---- vectors (DenseVector) ---------------------------------------------- */
int jfem_vec_create(jfem_ctx *ctx, long n, jfem_vec **out);
int jfem_vec_destroy(jfem_vec *v);
int jfem_vec_upload(jfem_vec *v, const double *host);
int jfem_vec_download(const jfem_vec *v, double *host);
int jfem_vec_fill(jfem_vec *v, double value);
/* DenseVector.inner (M/linalg/DenseVector.java:103-121) */
int jfem_vec_dot(const jfem_vec *x, const jfem_vec *y, double *out);
/* y = a*x + b*y */
int jfem_vec_axpby(double a, const jfem_vec *x, double b, jfem_vec *y);
/* dst = src, a device-to-device copy on the context stream (new DenseVector(Vector),
 * M/linalg/DenseVector.java:29-36); sizes must match */
int jfem_vec_copy(const jfem_vec *src, jfem_vec *dst);

/* ---- assembly: MatrixFESpace.evaluateCellIntegrals / writeCellIntegralsToMatrix
 *      (M/basic/MatrixFESpace.java:22-42,115-139) with TPCellIntegral
 *      (M/tensorproduct/TPCellIntegral.java:111-159) into a CSR whose pattern is
 *      the key set of SparseMatrix.getCoordinateEntryList()
 *      (M/linalg/SparseMatrix.java:402-418).  kinds/weights: nint integrals
 *      summed per entry (0 GRAD_GRAD, 1 VALUE_VALUE, 2 H1, 3 GRAD_VALUE,
 *      4 VALUE_GRAD); weights packed in kind order, one value per scalar-weighted
 *      integral and dim values (a constant vector) per GRAD_VALUE / VALUE_GRAD,
 *      which need a Gauss rule; cell_coeff (nullable, ncells host values)
 *      multiplies every integral's weight on that cell. ----------------------- */
int jfem_assemble(jfem_space *s, int nint, const int *kinds, const double *weights, const double *cell_coeff,
                  int quad_rule, jfem_mat **out);
/* TPRightHandSideIntegral VALUE (M/tensorproduct/TPRightHandSideIntegral.java:22-31)
 * of f(x) = sum_j amp[j] prod_a sin(kw[j*dim+a] pi x_a), into rhs (overwritten) */
int jfem_assemble_rhs_sines(jfem_space *s, int nterms, const int *kw, const double *amp, int quad_rule,
                            jfem_vec *rhs);
/* the same with a cosine factor on axis a of term j where cosine[j*dim+a] != 0
 * (cosine nullable = all sines): f(x) = sum_j amp[j] prod_a trig(kw[j*dim+a] pi x_a) */
int jfem_assemble_rhs_trig(jfem_space *s, int nterms, const int *kw, const int *cosine, const double *amp,
                           int quad_rule, jfem_vec *rhs);
/* rank of every DOF's node in CoordinateComparator order (M/linalg/CoordinateComparator.java:15-27;
 * first coordinate most significant), the order the mixed spaces number nodes in */
int jfem_space_lex_rank(jfem_space *s, int *out);

/* ---- Taylor-Hood Q(k+1)-Qk mixed system (M/mixed/TaylorHoodSpace.java:17-92) ----
 * Unknowns in TreeSet order of QkQkFunction.compareTo (M/mixed/QkQkFunction.java:82-93):
 * velocity component 0, ..., component dim-1, then pressure, each by node point
 * in CoordinateComparator order; size dim*ndofs(velocity) + ndofs(pressure).
 * `velocity` is the scalar Q(k+1) space and `pressure` the Qk space on the same grid. */
#define JFEM_MIXED_VELOCITY_GRAD_GRAD 0 /* MixedCellIntegral.fromVelocityIntegral(TPVectorCellIntegral GRAD_GRAD)
                                           (M/mixed/MixedCellIntegral.java:58-68, M/tensorproduct/TPVectorCellIntegral.java:80-88) */
#define JFEM_MIXED_DIV_VALUE 1          /* MixedTPCellIntegral DIV_VALUE (M/mixed/MixedTPCellIntegral.java:37-59),
                                           placed at (velocity, pressure) and (pressure, velocity) */
#define JFEM_MIXED_VALUE_GRAD 2         /* MixedTPCellIntegral VALUE_GRAD (M/mixed/MixedTPCellIntegral.java:50-57),
                                           velocity value . pressure gradient, placed like DIV_VALUE */
/* TaylorHoodSpace.writeCellIntegralsToMatrix (M/basic/MatrixFESpace.java:22-42 with
 * MixedCellIntegral.evaluateCellIntegral, M/mixed/MixedCellIntegral.java:93-124): at most one
 * integral of each kind, constant weights; pattern = COO key set as for jfem_assemble */
int jfem_th_assemble(jfem_space *velocity, jfem_space *pressure, int nint, const int *kinds, const double *weights,
                     int quad_rule, jfem_mat **out);
/* MixedRightHandSideIntegral.fromVelocityIntegral(TPVectorRightHandSideIntegral VALUE)
 * (M/mixed/MixedRightHandSideIntegral.java:35-74, M/tensorproduct/TPVectorRightHandSideIntegral.java:25-33)
 * for one component of the body force, f_c as in jfem_assemble_rhs_trig: ADDED into the
 * component's slice of the mixed vector rhs (writeCellIntegralsToRhs adds) */
int jfem_th_assemble_rhs_trig(jfem_space *velocity, jfem_space *pressure, int component, int nterms, const int *kw,
                              const int *cosine, const double *amp, int quad_rule, jfem_vec *rhs);

/* ---- matrices (SparseMatrix / VectorMultiplyable) ----------------------- */
int jfem_mat_from_csr(jfem_ctx *ctx, long rows, long cols, const long *rowptr, const int *colidx,
                      const double *vals, jfem_mat **out);
int jfem_mat_destroy(jfem_mat *A);
int jfem_mat_info(const jfem_mat *A, long *rows, long *cols, long *nnz);
int jfem_mat_export_csr(const jfem_mat *A, long *rowptr, int *colidx, double *vals);
/* setBoundaryValues -> overWriteValue (M/basic/AcceptsMatrixBoundaryValues.java:18-44):
 * rhs_j -= A_ji g_i for free j, row/column i zeroed (entries kept), A_ii = 1,
 * rhs_i = g_i.  values nullable (zero Dirichlet data). */
int jfem_apply_dirichlet(jfem_mat *A, jfem_vec *rhs, long nb, const int *nodes, const double *values);
/* VectorMultiplyable.mvMul (SparseMatrix.mvMul, M/linalg/SparseMatrix.java:479-489) */
int jfem_spmv(jfem_mat *A, const jfem_vec *x, jfem_vec *y);
/* VectorMultiplyable.tvMul (SparseMatrix.tvMul, M/linalg/SparseMatrix.java:491-501): y = A^T x.
 * The transpose is built on the first call (stable sort by column) and kept until A's values
 * change (jfem_apply_dirichlet drops it). */
int jfem_spmv_t(jfem_mat *A, const jfem_vec *x, jfem_vec *y);
/* y = A x with FP32 matrix values and FP32 vectors (device pointers to float) */
int jfem_spmv_f32(jfem_mat *A, const float *x, float *y);
/* repeat SpMV `reps` times on the context stream and return the mean kernel time (ms) */
int jfem_spmv_bench(jfem_mat *A, const jfem_vec *x, jfem_vec *y, int reps, int fp32, double *ms);

/* ---- Krylov solvers (IterativeSolver, M/linalg/IterativeSolver.java) ---- */
typedef struct {
    int iterations;   /* IterativeSolver.iterations */
    int fell_back;    /* BiCGStab breakdown fell back to (P)GMRES */
    double resnorm;   /* final ||b - A x||_2 (true residual) */
    double seconds;   /* wall time of the solve */
} jfem_solve_info;

/* solveCG / solvePCG (:48-162): stop when ||r||_2 <= tol (absolute), at most
 * min(n, maxit) iterations; x holds the start iterate on entry (the reference
 * default is e_0) and the solution on return.  M nullable (NULL = CG).
 * res_hist (nullable, maxit entries) receives ||r|| per iteration.
 * This and the other Krylov solvers apply A as operator.mvMul does (:134,141);
 * when A is the finest system of a hierarchy with a matrix-free finest level
 * (jfem_mg_level_matrix, values unchanged since) the product runs through that
 * operator, equal to the CSR product to rounding. */
int jfem_solve_cg(jfem_mat *A, jfem_mg *M, const jfem_vec *b, jfem_vec *x, double tol, long maxit,
                  jfem_solve_info *info, double *res_hist);
/* solveBiCGStab / solvePBiCGStab (:234-402) */
int jfem_solve_bicgstab(jfem_mat *A, jfem_mg *M, const jfem_vec *b, jfem_vec *x, double tol, long maxit,
                        jfem_solve_info *info);
/* GMRES.solve (M/linalg/GMRES.java:40-176): restarted, left preconditioned,
 * classical Gram-Schmidt, Givens; restart = ITERATIONS_BEFORE_RESTART (150) */
int jfem_solve_gmres(jfem_mat *A, jfem_mg *M, const jfem_vec *b, jfem_vec *x, double tol, int restart,
                     int max_cycles, jfem_solve_info *info);
/* Smoother.smooth(operator, rhs, iterate) (M/multigrid/Smoother.java:15) on any square matrix,
 * x in/out: JFEM_SM_JACOBI runs x { x += (D^-1 (b - A x)) omega }, |D_ii| < 1e-10 taken as 1
 * (JacobiSmoother.java:12-50); JFEM_SM_RICHARDSON runs x { r = b - A x; stop when ||r|| < 1e-10;
 * x += r omega } (RichardsonSmoother.java:13-50); JFEM_SM_CHEBYSHEV runs x a Chebyshev-Jacobi
 * polynomial of the given degree on [lmin, lmax] of D^-1 A (new). */
int jfem_smooth(jfem_mat *A, int kind, int runs, double omega, int degree, double lmin, double lmax,
                const jfem_vec *b, jfem_vec *x);
/* jfem_smooth without any device allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, kernels and results as jfem_smooth; its four work vectors
 * (A x, the residual, D^-1, the Chebyshev direction) live in a static device array of
 * 4 * JFEM_NOALLOC_MAX_ROWS doubles (32 MiB) that each GPU the library runs on holds until the
 * process ends.  A with more than JFEM_NOALLOC_MAX_ROWS rows is refused (JFEM_ERR_ARG).  All
 * calls on one GPU share that array: calls that overlap on one GPU (other threads, other
 * contexts' streams) race. */
#define JFEM_NOALLOC_MAX_ROWS 1048576
int jfem_smooth_noalloc(jfem_mat *A, int kind, int runs, double omega, int degree, double lmin, double lmax,
                        const jfem_vec *b, jfem_vec *x);
/* GMRES2.solve (M/linalg/GMRES2.java:34-104, Arnoldi M/linalg/Arnoldi.java:13-82): minimal
 * residual steps along M A-orthonormalised directions of the power sequence of M A, stop when
 * the recursively updated ||r||_2 <= tol (absolute), restart on a vanishing direction; M
 * nullable (identity).  x in: start iterate (reference default e_0).  maxit bounds the
 * iterations (unbounded in the reference); at most 512 directions are held between restarts. */
int jfem_solve_gmres2(jfem_mat *A, jfem_mg *M, const jfem_vec *b, jfem_vec *x, double tol, long maxit,
                      jfem_solve_info *info);
/* flexible CG (Notay 2000, Polak-Ribiere beta) in FP64 around a multigrid
 * preconditioner of any precision; not in the reference. */
int jfem_solve_fcg(jfem_mat *A, jfem_mg *M, const jfem_vec *b, jfem_vec *x, double tol, long maxit,
                   jfem_solve_info *info, double *res_hist);
/* Stokes saddle-point system of jfem_th_assemble (after Dirichlet data and the pressure
 * pin): right-preconditioned flexible GMRES(restart) with the block lower-triangular
 * preconditioner  z_u = M^-1 r_u (velocity_mg once per component, in the scalar space's
 * DOF order),  z_p = S^-1 (B z_u - r_p)  with B the pressure rows of A and the Schur
 * complement S ~ M_p / nu.  pressure_scale: ndofs(pressure) entries in the mixed order,
 * 1 / (row sum of M_p / nu) (the lumped mass), and -1 on a pinned (identity) pressure row.
 * S^-1 is the lumped inverse, or with opts->pressure_mass (M_p / nu on the scalar
 * pressure space, DOF order) mass_its Chebyshev steps preconditioned by the lumped mass
 * on [mass_lmin, mass_lmax] (Q1: 3^-dim, 1).  Stops when ||b - A x||_2 <= rtol ||b||_2 (the
 * Arnoldi estimate, confirmed on the true residual at each restart) or after opts->maxit
 * iterations.  The reference solves these systems with GMRES2 around a monolithic
 * multigrid (Examples/src/StokesMG2.java:26-129); this preconditioner is not in the
 * reference.  res_hist (nullable, maxit entries): relative residual estimate per iteration. */
typedef struct {
    int restart;                   /* FGMRES restart length */
    long maxit;                    /* iteration limit */
    jfem_mat *pressure_mass;       /* nullable: M_p / nu (scalar pressure space, DOF order) */
    int mass_its;                  /* Chebyshev steps on it (0 = lumped inverse only) */
    double mass_lmin, mass_lmax;   /* spectrum of (lumped M_p)^-1 M_p */
    jfem_mg *component_mg[2];      /* nullable: hierarchies equal to velocity_mg for components 1 and 2,
                                      built on other contexts; their V-cycles then run on those contexts'
                                      streams, concurrently with component 0's (Richardson smoothers
                                      synchronise with the host and gain nothing) */
} jfem_stokes_opts;
int jfem_solve_stokes_fgmres(jfem_mat *A, jfem_mg *velocity_mg, const jfem_vec *pressure_scale,
                             const jfem_stokes_opts *opts, const jfem_vec *b, jfem_vec *x, double rtol,
                             jfem_solve_info *info, double *res_hist);

/* ---- geometric multigrid (MGPreconditionerInterface / MGPreconditionerSpace,
 *      M/multigrid/MGPreconditionerInterface.java:125-379) -------------------- */
typedef struct {
    int smoother;      /* JFEM_SM_* */
    int runs;          /* Richardson/Jacobi runs, or Chebyshev polynomial applications */
    double omega;      /* Richardson/Jacobi relaxation */
    int degree;        /* Chebyshev degree */
    double cheb_ratio; /* Chebyshev interval [lmax/ratio, lmax] on D^-1 A */
    double cheb_safety;/* lmax = safety * (power-iteration estimate) */
    int power_its;     /* power iterations for the estimate */
    int cycles;        /* V-cycles per application (reference: 3) */
    int fmg;           /* 1: reference mvMul (copy BC, full-V correction, cycles V);
                        * 2: the same from v itself instead of its boundary values */
    int precision;     /* JFEM_FP64 or JFEM_FP32 */
    int galerkin;      /* 1: coarse systems A_{l-1} = P^T A_l P from the finest assembled one
                        * (AMGFromMatrix, M/multigrid/MGPreconditionerInterface.java:387-410)
                        * instead of re-assembly per level; the transfers keep their zero
                        * boundary conditions */
    int raw_transfers; /* 1: MGSpace (M/multigrid/MGSpace.java:25-206): restriction and
                        * prolongation by the prolongation matrix without zero boundary values.
                        * With fmg = 2 and cycles = 1 the application is MGPReconditioner2.mvMul
                        * (M/multigrid/MGPReconditioner2.java:38-76): v + fullVCycleSolver(v - A v) */
    int matrix_free;   /* 1: smoothing and residuals on re-assembled levels (every level, or the
                        * finest of a Galerkin hierarchy) apply the level operator matrix-free: the
                        * cell element matrices of GRAD_GRAD / VALUE_VALUE / H1 by 1D sum
                        * factorisation, summed per node, on 2D / 3D grids of degree 1 or 2 (others
                        * keep the CSR product); 2: the same on 3D grids only (2D rows hold 9 or 25
                        * entries, where the CSR product is as cheap); 0: CSR everywhere.  The level
                        * matrices are still assembled (jfem_mg_level_matrix); the products agree
                        * with them to rounding.  New; the reference's smoothers call
                        * getSystem(level).mvMul (M/multigrid/MGPreconditionerInterface.java:125-198). */
} jfem_mg_opts;

/* levels: nlev spaces, coarsest first, each with twice the cells of the
 * previous one; every level is assembled with the same integrals and zero
 * Dirichlet data on the whole boundary (createSystem per level).  cell_coeff
 * (nullable) is the finest level's per-cell coefficient; coarser levels use
 * the mean of their 2^dim children. */
int jfem_mg_create(jfem_space *const *spaces, int nlev, int nint, const int *kinds, const double *weights,
                   const double *cell_coeff, int quad_rule, const jfem_mg_opts *opts, jfem_mg **out);
int jfem_mg_destroy(jfem_mg *mg);
/* the finest level's (Dirichlet-eliminated) system matrix, owned by the hierarchy */
int jfem_mg_level_matrix(jfem_mg *mg, int level, jfem_mat **out);
/* getSystem(level).mvMul(x) (MGPreconditionerInterface.java:158): y = A_level x with the operator
 * the hierarchy's cycles use (matrix-free where jfem_mg_level_matrix_free reports 1), FP64 */
int jfem_mg_level_apply(jfem_mg *mg, int level, const jfem_vec *x, jfem_vec *y);
/* whether the hierarchy applies level `level` matrix-free (out: 1) or by its CSR matrix (0); new */
int jfem_mg_level_matrix_free(jfem_mg *mg, int level, int *out);
/* repeat one Chebyshev step of level `level` (the smoothers' launch: matrix-free where
 * jfem_mg_level_matrix_free reports 1, else through the CSR matrix) `reps` times on the context
 * stream from x into y and return the mean kernel time (ms), as jfem_spmv_bench does for the
 * SpMV; b and the direction are the level's own work vectors; FP64 hierarchies; new */
int jfem_mg_level_bench(jfem_mg *mg, int level, const jfem_vec *x, jfem_vec *y, int reps, double *ms);
/* MGPreconditionerInterface.mvMul: z = M r */
int jfem_mg_apply(jfem_mg *mg, const jfem_vec *r, jfem_vec *z);
/* one V-cycle mgStepV(finest, x, b) updating x in place */
int jfem_mg_vcycle(jfem_mg *mg, jfem_vec *x, const jfem_vec *b);
/* wCycle -> mgStepW (MGPreconditionerInterface.java:200-305): pre-smoothing, two coarse
 * corrections each by a V-cycle on the next coarser level, post-smoothing; x in/out */
int jfem_mg_wcycle(jfem_mg *mg, jfem_vec *x, const jfem_vec *b);
/* Chebyshev interval upper bound used on each level (after safety factor) */
int jfem_mg_lmax(jfem_mg *mg, double *out);
/* prolongationOperators.get(level).mvMul / .tvMul
 * (MGPreconditionerSpace.java:20,61-70): fine = P coarse and coarse = P^T fine
 * between level `level` and level + 1, FP64, no boundary treatment */
int jfem_mg_prolongate(jfem_mg *mg, int level, const jfem_vec *coarse, jfem_vec *fine);
int jfem_mg_restrict(jfem_mg *mg, int level, const jfem_vec *fine, jfem_vec *coarse);

/* ---- learned smoother network (not in the reference) --------------------
 * four 3^dim-tap convolutions on the node lattice: cin -> 32 -> 32 -> 32 -> 1
 * with ReLU; BF16 MFMA for the 32->32 layers, FP32 accumulation, FP32 master
 * weights and Adam.  Lattice arrays are in lattice order, axis 0 fastest;
 * inputs are pixel-major (np x cin). */
int jfem_nn_create(jfem_ctx *ctx, int dim, int cin, unsigned long long seed, jfem_nn **out);
int jfem_nn_destroy(jfem_nn *nn);
long jfem_nn_num_params(const jfem_nn *nn);
/* parameters in the order W0[t][32][cin], b0[32], W1[t][32][32], b1, W2, b2, W3[t][32], b3 */
int jfem_nn_get_params(jfem_nn *nn, float *out);
int jfem_nn_set_params(jfem_nn *nn, const float *in);
int jfem_nn_forward_host(jfem_nn *nn, const int *shape, const float *in, float *out);
/* gradient of sum(gout * forward(in)) with respect to the parameters */
int jfem_nn_gradient_host(jfem_nn *nn, const int *shape, const float *in, const float *gout, float *grad);
/* one Adam step (grad NULL: the gradient left by the last backward pass) */
int jfem_nn_adam_step(jfem_nn *nn, const float *grad, float lr);
/* learned smoother on the finest multigrid level: with r = b - A x and
 * u = N(r / rms(r)), step_rule 0 takes x += rms(r) u, 1 the minimal-residual
 * step x += alpha u with alpha = (r.Au)/(Au.Au), 2 the minimal-energy step
 * alpha = (u.r)/(u.Au) (never increases the A-norm of the error); `runs`
 * steps per smoothing call.  Input channels: the
 * normalised residual r^, then r^ times the mean log-coefficient around each
 * node (cell_log_coeff of the finest grid, nullable = 0) and, for cin = 2 + dim,
 * r^ times the lattice parity per axis.  Attaching a network sets its biases
 * to zero and keeps them there, so N(0) = 0 and N(c r) = c N(r) for c > 0.  The hierarchy's Chebyshev smoother is the
 * fallback, taken once a cycle (or a flexible-CG iteration) reduces the
 * residual by less than fallback_ratio.  nn NULL detaches. */
int jfem_mg_set_learned(jfem_mg *mg, jfem_nn *nn, const double *cell_log_coeff, int runs, double fallback_ratio,
                        int step_rule);
int jfem_mg_set_learned_active(jfem_mg *mg, int active);
/* which smoothing calls of a V-cycle on the finest level use the attached
 * network: 1 pre-smoothing, 2 post-smoothing, 3 both (default); the other
 * side keeps the hierarchy's smoother.  New; the reference's
 * MGPreconditionerSpace takes one smoother per level for both sides
 * (M/multigrid/MGPreconditionerInterface.java:125-198). */
int jfem_mg_set_learned_sides(jfem_mg *mg, int sides);
int jfem_mg_learned_status(jfem_mg *mg, int *active, int *fell_back);
typedef struct {
    int rhs_per_field;        /* seeded random right-hand sides */
    int cycles_per_rhs;       /* (residual, error) samples per right-hand side: one per plain V-cycle */
    double lr;                /* Adam learning rate */
    unsigned long long seed;
    int mode;                 /* samples: 0 = residuals of plain V-cycles, 1 = A e for white-noise
                                 errors e, 2 = alternate 0 and 1, 3 = alternate the residual at cycle
                                 entry and the residual before finest post-smoothing;
                                 4 = as 3 on A x = A x* for random x* (the error is known) with the
                                 two-grid loss below; 5 = as 4, but every second right-hand side
                                 advances with the V-cycle that uses the learned smoother */
} jfem_train_opts;
/* trains the network attached to mg on (residual, error) pairs of its plain
 * V-cycles, one Adam step per sample.  Modes 0-3: the loss is the energy of the
 * error left by the learned step, |e - N(r)|_A^2 - |e|_A^2 per node (input
 * normalised); mode 4: the energy of the error left by the learned step followed
 * by the coarse-grid correction, |T (e - N(r))|_A^2, T = I - P A_c^-1 P^T A, and
 * loss_hist gets |T (e - N(r))|_A^2 / |e|_A^2 per step.  loss_hist is nullable. */
int jfem_nn_train_mg(jfem_mg *mg, const jfem_train_opts *opts, double *loss_hist, int *nsteps);
/* stationary multigrid iteration to ||b - A x|| <= tol (x: start iterate in, solution out) */
int jfem_mg_solve(jfem_mg *mg, const jfem_vec *b, jfem_vec *x, double tol, int maxcycles, jfem_solve_info *info,
                  double *res_hist);
/* versioned checkpoint of weights and optimiser state */
int jfem_nn_save(jfem_nn *nn, const char *path);
int jfem_nn_load(jfem_ctx *ctx, const char *path, jfem_nn **out);

#ifdef __cplusplus
}
#endif

#endif
