/*
 * jfem_oracle.c -- CPU restatement of the JFEM (Tovermodus/JFEM, Java) solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or the timed CPU baseline.  The product (libjfem_hip.so)
 * never links, calls or falls back to it.
 *
 * Single-threaded, compiled with -O2 -ffp-contract=off (no FMA contraction, no
 * fast-math) so every floating-point operation happens in the order the Java
 * source performs it.  Each function cites the reference file:line it restates
 * (paths relative to the reference repository; M/ = Core/src/main/java/).
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   quadrature / Lagrange basis  : closed forms, exactness of Gauss-n on degree 2n-1
 *   DOF numbering (orc_dof_table): brute-force first-appearance replay vs closed form
 *   cell integrals               : golden arrays of ContinuousIntegralTest.java:23-40,
 *                                  exact 1D K/M tensor-product closed forms
 *   assembly / Dirichlet / CSR   : symmetry, zero row sums, pattern closed form,
 *                                  scipy-built reference matrices
 *   CG/PCG/BiCGStab/GMRES        : BiCGStabTest/GMResTest known answers, scipy LU
 *   MG                           : prolongation reproduces linear functions,
 *                                  P^T A_f P == A_c, LaplaceMG iteration bound
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXD 3

/* ------------------------------------------------------------------------- */
/* QuadratureRule1D (M/tensorproduct/QuadratureRule1D.java:11-64, ctor :83-92) */
/* rule ids: 1..5 = Gauss1..Gauss5, 12..16 = GaussLobatto2..6                  */
/* ------------------------------------------------------------------------- */
int orc_quad(int rule, double *pts, double *wts)
{
    double p[6], w[6];
    int n = 0;
    switch (rule) {
    case 1: n = 1; w[0] = 2; p[0] = 0; break;
    case 2: n = 2; w[0] = 1; w[1] = 1; p[0] = -sqrt(1. / 3); p[1] = sqrt(1. / 3); break;
    case 3: n = 3; w[0] = 5. / 9; w[1] = 8. / 9; w[2] = 5. / 9;
        p[0] = -sqrt(3. / 5); p[1] = 0; p[2] = sqrt(3. / 5); break;
    case 4: n = 4;
        w[0] = (18. - sqrt(30.)) / 36; w[1] = (18. + sqrt(30.)) / 36;
        w[2] = (18. + sqrt(30.)) / 36; w[3] = (18. - sqrt(30.)) / 36;
        p[0] = -sqrt(3. / 7 + 2. / 7 * sqrt(6. / 5)); p[1] = -sqrt(3. / 7 - 2. / 7 * sqrt(6. / 5));
        p[2] = sqrt(3. / 7 - 2. / 7 * sqrt(6. / 5));  p[3] = sqrt(3. / 7 + 2. / 7 * sqrt(6. / 5));
        break;
    case 5: n = 5;
        w[0] = (322.0 - 13 * sqrt(70)) / 900; w[1] = (322.0 + 13 * sqrt(70)) / 900;
        w[2] = 128.0 / 225;
        w[3] = (322.0 + 13 * sqrt(70)) / 900; w[4] = (322.0 - 13 * sqrt(70)) / 900;
        p[0] = -sqrt(5.0 + 2 * sqrt(10.0 / 7)) / 3; p[1] = -sqrt(5.0 - 2 * sqrt(10.0 / 7)) / 3;
        p[2] = 0;
        p[3] = sqrt(5.0 - 2 * sqrt(10.0 / 7)) / 3;  p[4] = sqrt(5.0 + 2 * sqrt(10.0 / 7)) / 3;
        break;
    case 12: n = 2; w[0] = 1. / 2; w[1] = 1. / 2; p[0] = -1; p[1] = 1; break;
    case 13: n = 3; w[0] = 1. / 3; w[1] = 4. / 3; w[2] = 1. / 3; p[0] = -1; p[1] = 0; p[2] = 1; break;
    case 14: n = 4; w[0] = 1. / 6; w[1] = 5. / 6; w[2] = 5. / 6; w[3] = 1. / 6;
        p[0] = -1; p[1] = -sqrt(1. / 5); p[2] = sqrt(1. / 5); p[3] = 1; break;
    case 15: n = 5; w[0] = 1. / 10; w[1] = 49. / 90; w[2] = 32. / 45; w[3] = 49. / 90; w[4] = 1. / 10;
        p[0] = -1; p[1] = -sqrt(3. / 7); p[2] = 0; p[3] = sqrt(3. / 7); p[4] = 1; break;
    case 16: n = 6;
        w[0] = 1. / 15; w[1] = (14. - sqrt(7)) / 30; w[2] = (14. + sqrt(7)) / 30;
        w[3] = (14. + sqrt(7)) / 30; w[4] = (14. - sqrt(7)) / 30; w[5] = 1. / 15;
        p[0] = -1; p[1] = -sqrt(1. / 3 + 2 * sqrt(7) / 21); p[2] = -sqrt(1. / 3 - 2 * sqrt(7) / 21);
        p[3] = +sqrt(1. / 3 - 2 * sqrt(7) / 21); p[4] = +sqrt(1. / 3 + 2 * sqrt(7) / 21); p[5] = 1;
        break;
    default: return -1;
    }
    for (int i = 0; i < n; i++) {      /* QuadratureRule1D ctor :87-91 */
        pts[i] = p[i] / 2 + 0.5;
        wts[i] = w[i] / 2;
    }
    return n;
}

/* ------------------------------------------------------------------------- */
/* LagrangeBasisFunction1D (M/tensorproduct/LagrangeBasisFunction1D.java:55-154) */
/* ------------------------------------------------------------------------- */
double orc_lagrange_value(double pos, int l, int k)
{
    switch (k) {
    case 1:
        if (l == 0) return 1.0 - pos;
        if (l == 1) return pos;
        return 0;
    case 2:
        if (l == 0) return 1.0 - 3.0 * pos + 2.0 * pos * pos;
        if (l == 1) return (pos - pos * pos) * 4.0;
        if (l == 2) return 2.0 * pos * pos - pos;
        return 0;
    default: {
        double ret = 1;
        for (int i = 0; i <= k; i++)
            if (i != l)
                ret *= (pos - 1. * i / k) / (1. * l / k - 1. * i / k);
        return ret;
    }
    }
}

double orc_lagrange_deriv(double pos, int l, int k)
{
    switch (k) {
    case 1:
        if (l == 0) return -1.0;
        if (l == 1) return 1.0;
        return 0;
    case 2:
        if (l == 0) return -3.0 + 4.0 * pos;
        if (l == 1) return (1.0 - 2.0 * pos) * 4.0;
        if (l == 2) return 4.0 * pos - 1.0;
        return 0;
    default: {
        double derret = 0;
        for (int j = 0; j <= k; j++) {
            if (j == l) continue;
            double ret = 1. / (1. * l / k - 1. * j / k);
            for (int i = 0; i <= k; i++)
                if (i != l && i != j)
                    ret *= (pos - 1. * i / k) / (1. * l / k - 1. * i / k);
            derret += ret;
        }
        return derret;
    }
    }
}

/* Cell1D (M/tensorproduct/geometry/Cell1D.java:45-69); tolerance from
 * PerformanceArguments.doubleTolerance = 1e-10 (M/basic/PerformanceArguments.java). */
static const double ORC_TOL = 1e-10;
static int cell1d_in(double s, double e, double pos)
{
    return pos >= s - 100 * ORC_TOL && pos <= e + 100 * ORC_TOL;
}
/* LagrangeBasisFunction1D.value/derivative (:161-176) */
static double lag_val_cell(double s, double e, int l, int k, double pos)
{
    if (cell1d_in(s, e, pos)) return orc_lagrange_value((pos - s) / (e - s), l, k);
    return 0;
}
static double lag_der_cell(double s, double e, int l, int k, double pos)
{
    if (cell1d_in(s, e, pos)) return orc_lagrange_deriv((pos - s) / (e - s), l, k) / (e - s);
    return 0;
}

/* CartesianGrid.createCells1D (M/tensorproduct/geometry/CartesianGrid.java:26-49) */
void orc_cells1d(double start, double end, int n, double *s, double *e)
{
    const double width = (end - start) / n;
    for (int j = 0; j < n; j++) {
        s[j] = j * width + start;
        e[j] = (j + 1) * width + start;
    }
}

/* ------------------------------------------------------------------------- */
/* DOF numbering, brute-force replay of ContinuousTPFESpace.assembleFunctions  */
/* (M/tensorproduct/ContinuousTPFESpace.java:29-50): cells in IntCoordinates   */
/* order (last axis fastest, CartesianGrid.java:142-158), local index with      */
/* axis 0 fastest (ContinuousTPShapeFunction.decomposeIndex :64-73), a node is  */
/* new the first time its point appears.  Node identity is taken on the node    */
/* lattice, which coincides with the reference's 1e-10 coordinate comparison on */
/* any non-degenerate grid.  cell_dofs[cell*(k+1)^d + local] = global index.    */
/* ------------------------------------------------------------------------- */
long orc_dof_table(int d, int k, const int *cells, int *cell_dofs)
{
    long nodes_per_axis[ORC_MAXD], nl = 1, nc = 1;
    for (int a = 0; a < d; a++) { nodes_per_axis[a] = (long)k * cells[a] + 1; nl *= nodes_per_axis[a]; nc *= cells[a]; }
    int *seen = (int *)malloc(sizeof(int) * nl);
    for (long i = 0; i < nl; i++) seen[i] = -1;
    int nloc = 1;
    for (int a = 0; a < d; a++) nloc *= (k + 1);
    long next = 0;
    for (long c = 0; c < nc; c++) {
        int cc[ORC_MAXD];
        long t = c;
        for (int a = d - 1; a >= 0; a--) { cc[a] = (int)(t % cells[a]); t /= cells[a]; }
        for (int li = 0; li < nloc; li++) {
            int r = li;
            long lat = 0;                    /* lattice linear index, axis 0 slowest */
            int lat_a[ORC_MAXD];
            for (int a = 0; a < d; a++) { lat_a[a] = cc[a] * k + r % (k + 1); r /= (k + 1); }
            for (int a = 0; a < d; a++) lat = lat * nodes_per_axis[a] + lat_a[a];
            if (seen[lat] < 0) seen[lat] = (int)next++;
            cell_dofs[c * nloc + li] = seen[lat];
        }
    }
    free(seen);
    return next;
}

/* ------------------------------------------------------------------------- */
/* Cell integrals: TPCellIntegral.evaluateCellIntegral (M/tensorproduct/      */
/* TPCellIntegral.java:111-159) over integrateNonTensorProduct (:78-102), with  */
/* ContinuousTPShapeFunction.fastValueInCell/fastGradientInCell (:154-198) and  */
/* DenseVector.inner (M/linalg/DenseVector.java:103-121).                       */
/* ------------------------------------------------------------------------- */
enum { ORC_GRAD_GRAD = 0, ORC_VALUE_VALUE = 1, ORC_H1 = 2, ORC_GRAD_VALUE = 3, ORC_VALUE_GRAD = 4 };

typedef double (*orc_sfun)(const double *x, int d, void *ctx);
typedef void (*orc_vfun)(const double *x, int d, double *out, void *ctx);

typedef struct {
    int d, k;
    double s[ORC_MAXD], e[ORC_MAXD];  /* component cells */
} orc_cell;

static double sf_value(const orc_cell *K, const int *l, const double *x)
{
    double ret = 1;
    for (int i = 0; i < K->d; i++) ret *= lag_val_cell(K->s[i], K->e[i], l[i], K->k, x[i]);
    return ret;
}
static void sf_grad(const orc_cell *K, const int *l, const double *x, double *g)
{
    for (int i = 0; i < K->d; i++) {
        double comp = 1;
        for (int j = 0; j < K->d; j++) {
            if (i == j) comp *= lag_der_cell(K->s[j], K->e[j], l[j], K->k, x[j]);
            else comp *= lag_val_cell(K->s[j], K->e[j], l[j], K->k, x[j]);
        }
        g[i] = comp;
    }
}
static double inner(const double *a, const double *b, int n)
{
    double r = 0;
    for (int i = 0; i < n; i++) r += a[i] * b[i];
    return r;
}

/* Per-cell 1D tables: the values the reference's basis evaluations produce at
 * the cell's quadrature points (computed once per cell; the integral below
 * consumes exactly these numbers in the reference's order). */
typedef struct {
    int Q;
    double pw[ORC_MAXD][2][8];
    double val[ORC_MAXD][8][8];   /* [axis][l][q] */
    double der[ORC_MAXD][8][8];
} orc_tabs;

static void cell_tabs(const orc_cell *K, int rule, orc_tabs *T)
{
    double qp[8], qw[8];
    T->Q = orc_quad(rule, qp, qw);
    for (int j = 0; j < K->d; j++)                /* Cell1D.distributeQuadrature :24-33 */
        for (int i = 0; i < T->Q; i++) {
            T->pw[j][0][i] = qp[i] * (K->e[j] - K->s[j]) + K->s[j];
            T->pw[j][1][i] = qw[i] * (K->e[j] - K->s[j]);
            for (int l = 0; l <= K->k; l++) {
                T->val[j][l][i] = lag_val_cell(K->s[j], K->e[j], l, K->k, T->pw[j][0][i]);
                T->der[j][l][i] = lag_der_cell(K->s[j], K->e[j], l, K->k, T->pw[j][0][i]);
            }
        }
}

/* One cell integral between local functions u (shapeFunction1) and v
 * (shapeFunction2), local multi-indices lu/lv.  Weight: wfun(x) if given, else
 * the constant wconst; vector weight vfun for GRAD_VALUE/VALUE_GRAD. */
static double integral_tabs(const orc_cell *K, const orc_tabs *T, int kind, const int *lu, const int *lv,
                            double wconst, orc_sfun wfun, orc_vfun vfun, void *ctx)
{
    const int d = K->d, Q = T->Q;
    int nq = 1;
    for (int j = 0; j < d; j++) nq *= Q;
    double ret = 0, x[ORC_MAXD];
    for (int q = 0; q < nq; q++) {
        int c[ORC_MAXD], t = q;
        for (int j = d - 1; j >= 0; j--) { c[j] = t % Q; t /= Q; }   /* last axis fastest */
        for (int j = 0; j < d; j++) x[j] = T->pw[j][0][c[j]];
        double val, g1[ORC_MAXD], g2[ORC_MAXD], wv[ORC_MAXD];
        double v1 = 1, v2 = 1;                   /* fastValueInCell :154-171 */
        for (int i = 0; i < d; i++) { v1 *= T->val[i][lu[i]][c[i]]; v2 *= T->val[i][lv[i]][c[i]]; }
        for (int i = 0; i < d; i++) {            /* fastGradientInCell :173-198 */
            double c1 = 1, c2 = 1;
            for (int j = 0; j < d; j++) {
                if (i == j) { c1 *= T->der[j][lu[j]][c[j]]; c2 *= T->der[j][lv[j]][c[j]]; }
                else { c1 *= T->val[j][lu[j]][c[j]]; c2 *= T->val[j][lv[j]][c[j]]; }
            }
            g1[i] = c1; g2[i] = c2;
        }
        switch (kind) {
        case ORC_GRAD_GRAD: val = inner(g1, g2, d) * (wfun ? wfun(x, d, ctx) : wconst); break;
        case ORC_VALUE_VALUE: val = v1 * v2 * (wfun ? wfun(x, d, ctx) : wconst); break;
        case ORC_H1: val = (v1 * v2 + inner(g1, g2, d)) * (wfun ? wfun(x, d, ctx) : wconst); break;
        case ORC_GRAD_VALUE: vfun(x, d, wv, ctx); val = inner(g1, wv, d) * v2; break;
        case ORC_VALUE_GRAD: vfun(x, d, wv, ctx); val = inner(g2, wv, d) * v1; break;
        default: return NAN;
        }
        for (int j = 0; j < d; j++) val *= T->pw[j][1][c[j]];
        ret += val;
    }
    return ret;
}

double orc_cell_integral_ex(const orc_cell *K, int kind, int rule, const int *lu, const int *lv,
                            double wconst, orc_sfun wfun, orc_vfun vfun, void *ctx)
{
    orc_tabs T;
    cell_tabs(K, rule, &T);
    return integral_tabs(K, &T, kind, lu, lv, wconst, wfun, vfun, ctx);
}

/* Convenience entry for ctypes: cell given by starts/ends arrays. */
double orc_cell_integral(int d, int k, const double *s, const double *e, int kind, int rule,
                         const int *lu, const int *lv, double wconst, orc_sfun wfun, orc_vfun vfun, void *ctx)
{
    orc_cell K;
    K.d = d; K.k = k;
    for (int a = 0; a < d; a++) { K.s[a] = s[a]; K.e[a] = e[a]; }
    return orc_cell_integral_ex(&K, kind, rule, lu, lv, wconst, wfun, vfun, ctx);
}

/* RHS VALUE integral, TPRightHandSideIntegral.evaluateRightHandSideIntegral
 * (M/tensorproduct/TPRightHandSideIntegral.java:22-31): phi_v(x) * f(x). */
static double cell_rhs_integral(const orc_cell *K, int rule, const int *lv, orc_sfun f, void *ctx)
{
    double qp[8], qw[8];
    const int Q = orc_quad(rule, qp, qw);
    const int d = K->d;
    double pw[ORC_MAXD][2][8];
    for (int j = 0; j < d; j++)
        for (int i = 0; i < Q; i++) {
            pw[j][0][i] = qp[i] * (K->e[j] - K->s[j]) + K->s[j];
            pw[j][1][i] = qw[i] * (K->e[j] - K->s[j]);
        }
    int nq = 1;
    for (int j = 0; j < d; j++) nq *= Q;
    double ret = 0, x[ORC_MAXD];
    for (int q = 0; q < nq; q++) {
        int c[ORC_MAXD], t = q;
        for (int j = d - 1; j >= 0; j--) { c[j] = t % Q; t /= Q; }
        for (int j = 0; j < d; j++) x[j] = pw[j][0][c[j]];
        double val = sf_value(K, lv, x) * f(x, d, ctx);
        for (int j = 0; j < d; j++) val *= pw[j][1][c[j]];
        ret += val;
    }
    return ret;
}

/* ------------------------------------------------------------------------- */
/* Assembly: MatrixFESpace.writeCellIntegralsToMatrix / loopMatrixViaCell /    */
/* addToMatrix (M/basic/MatrixFESpace.java:29-42,115-139) into SparseMatrix    */
/* COO semantics (M/linalg/SparseMatrix.java:84-102 dummy (0,0) entry, add     */
/* skips exact zeros :273-294, getCoordinateEntryList sums duplicates and keeps */
/* explicit zeros :402-418), emitted as canonical CSR (rows ascending, columns */
/* ascending).  Duplicates are summed in append order (cells ascending).        */
/* ------------------------------------------------------------------------- */
typedef struct {
    long n, nnz;
    long *rowptr;   /* n+1 */
    int *cols;      /* nnz */
    double *vals;   /* nnz */
} orc_csr;

void orc_csr_free(orc_csr *A)
{
    if (!A) return;
    free(A->rowptr); free(A->cols); free(A->vals); free(A);
}

typedef struct { int row, col; long seq; double val; } coo_t;

static int coo_cmp(const void *pa, const void *pb)
{
    const coo_t *a = (const coo_t *)pa, *b = (const coo_t *)pb;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->col != b->col) return a->col < b->col ? -1 : 1;
    return a->seq < b->seq ? -1 : (a->seq > b->seq);
}

static orc_csr *coo_to_csr(long n, coo_t *coo, long m)
{
    qsort(coo, m, sizeof(coo_t), coo_cmp);
    orc_csr *A = (orc_csr *)calloc(1, sizeof(orc_csr));
    A->n = n;
    A->rowptr = (long *)calloc(n + 1, sizeof(long));
    A->cols = (int *)malloc(sizeof(int) * (m > 0 ? m : 1));
    A->vals = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    long nnz = 0;
    for (long i = 0; i < m;) {
        long j = i;
        double acc = 0;                      /* getCoordinateEntryList: value += previous */
        while (j < m && coo[j].row == coo[i].row && coo[j].col == coo[i].col) { acc = coo[j].val + acc; j++; }
        A->cols[nnz] = coo[i].col;
        A->vals[nnz] = acc;
        A->rowptr[coo[i].row + 1]++;
        nnz++;
        i = j;
    }
    for (long r = 0; r < n; r++) A->rowptr[r + 1] += A->rowptr[r];
    A->nnz = nnz;
    return A;
}

typedef struct {
    int d, k, rule;
    int cells[ORC_MAXD];
    double start[ORC_MAXD], end[ORC_MAXD];
    int nint;
    const int *kinds;         /* nint integral kinds */
    const double *wconst;     /* nint constant weights */
    const double *cell_coeff; /* nullable: weight of every integral on cell c is cell_coeff[c]*wconst[i] */
} orc_asm_desc;

static void cell_geometry(const orc_asm_desc *D, long c, int *cc, orc_cell *K, double s1d[][4096], double e1d[][4096])
{
    long t = c;
    for (int a = D->d - 1; a >= 0; a--) { cc[a] = (int)(t % D->cells[a]); t /= D->cells[a]; }
    K->d = D->d; K->k = D->k;
    for (int a = 0; a < D->d; a++) { K->s[a] = s1d[a][cc[a]]; K->e[a] = e1d[a][cc[a]]; }
}

static double (*g_s1d)[4096];
static double (*g_e1d)[4096];

static int setup_grid(const orc_asm_desc *D)
{
    for (int a = 0; a < D->d; a++) if (D->cells[a] > 4096) return -1;
    g_s1d = (double (*)[4096])malloc(sizeof(double) * 4096 * ORC_MAXD);
    g_e1d = (double (*)[4096])malloc(sizeof(double) * 4096 * ORC_MAXD);
    for (int a = 0; a < D->d; a++) orc_cells1d(D->start[a], D->end[a], D->cells[a], g_s1d[a], g_e1d[a]);
    return 0;
}
static void teardown_grid(void) { free(g_s1d); free(g_e1d); g_s1d = NULL; g_e1d = NULL; }

/* constant vector weight for GRAD_VALUE / VALUE_GRAD: ctx holds d values (scaled per cell) */
static void const_vfun(const double *x, int d, double *out, void *ctx)
{
    (void)x;
    const double *w = (const double *)ctx;
    for (int i = 0; i < d; i++) out[i] = w[i];
}

orc_csr *orc_assemble_w(int d, int k, const int *cells, const double *start, const double *end, int nint,
                        const int *kinds, const double *wconst, const double *wvec, const double *cell_coeff,
                        int rule);

orc_csr *orc_assemble(int d, int k, const int *cells, const double *start, const double *end,
                      int nint, const int *kinds, const double *wconst, const double *cell_coeff, int rule)
{
    return orc_assemble_w(d, k, cells, start, end, nint, kinds, wconst, NULL, cell_coeff, rule);
}

/* as orc_assemble; wvec (nint x d, nullable) holds the constant vector weight of each
 * GRAD_VALUE / VALUE_GRAD integral, multiplied by the cell coefficient on each cell */
orc_csr *orc_assemble_w(int d, int k, const int *cells, const double *start, const double *end, int nint,
                        const int *kinds, const double *wconst, const double *wvec, const double *cell_coeff,
                        int rule)
{
    orc_asm_desc D;
    /* Every kind but GRAD_GRAD evaluates value(x) / gradient(x), i.e. in the first support cell
     * containing x (ScalarShapeFunction.java:26-39; TPCellIntegral.java:124-157); this restatement
     * evaluates in the integrated cell, which is the same at interior points only.  Gauss-Lobatto
     * points lie on cell faces, so those kinds are refused with them. */
    for (int i = 0; i < nint; i++)
        if (kinds[i] != ORC_GRAD_GRAD && !(rule >= 1 && rule <= 5)) return NULL;
    memset(&D, 0, sizeof(D));
    D.d = d; D.k = k; D.rule = rule; D.nint = nint; D.kinds = kinds; D.wconst = wconst; D.cell_coeff = cell_coeff;
    long nc = 1;
    int nloc = 1;
    for (int a = 0; a < d; a++) { D.cells[a] = cells[a]; D.start[a] = start[a]; D.end[a] = end[a]; nc *= cells[a]; nloc *= k + 1; }
    if (setup_grid(&D)) return NULL;
    int *cd = (int *)malloc(sizeof(int) * nc * nloc);
    const long n = orc_dof_table(d, k, cells, cd);
    long cap = nc * (long)nloc * nloc + 1;
    coo_t *coo = (coo_t *)malloc(sizeof(coo_t) * cap);
    long m = 0;
    coo[m].row = 0; coo[m].col = 0; coo[m].val = 0.0; coo[m].seq = m; m++;   /* dummy entry */
    int lu[ORC_MAXD], lv[ORC_MAXD];
    for (long c = 0; c < nc; c++) {
        int cc[ORC_MAXD];
        orc_cell K;
        cell_geometry(&D, c, cc, &K, g_s1d, g_e1d);
        orc_tabs T;
        cell_tabs(&K, rule, &T);
        for (int iu = 0; iu < nloc; iu++) {          /* function1 = u (column) */
            int r = iu;
            for (int a = 0; a < d; a++) { lu[a] = r % (k + 1); r /= (k + 1); }
            for (int iv = 0; iv < nloc; iv++) {      /* function2 = v (row) */
                r = iv;
                for (int a = 0; a < d; a++) { lv[a] = r % (k + 1); r /= (k + 1); }
                double integral = 0;
                for (int i = 0; i < nint; i++) {
                    const double w = cell_coeff ? cell_coeff[c] * wconst[i] : wconst[i];
                    if (kinds[i] == ORC_GRAD_VALUE || kinds[i] == ORC_VALUE_GRAD) {
                        double wv[ORC_MAXD];
                        for (int a = 0; a < d; a++)
                            wv[a] = cell_coeff ? cell_coeff[c] * wvec[i * d + a] : wvec[i * d + a];
                        integral += integral_tabs(&K, &T, kinds[i], lu, lv, 0.0, NULL, const_vfun, wv);
                    } else {
                        integral += integral_tabs(&K, &T, kinds[i], lu, lv, w, NULL, NULL, NULL);
                    }
                }
                if (integral != 0) {
                    coo[m].row = cd[c * nloc + iv];
                    coo[m].col = cd[c * nloc + iu];
                    coo[m].val = integral;
                    coo[m].seq = m;
                    m++;
                }
            }
        }
    }
    orc_csr *A = coo_to_csr(n, coo, m);
    free(coo); free(cd);
    teardown_grid();
    return A;
}

/* Element matrix of one cell of a grid (for fixtures and the class-dedup check). */
int orc_element_matrix(int d, int k, const int *cells, const double *start, const double *end, long cell,
                       int nint, const int *kinds, const double *wconst, int rule, double *out)
{
    orc_asm_desc D;
    memset(&D, 0, sizeof(D));
    D.d = d; D.k = k;
    int nloc = 1;
    for (int a = 0; a < d; a++) { D.cells[a] = cells[a]; D.start[a] = start[a]; D.end[a] = end[a]; nloc *= k + 1; }
    if (setup_grid(&D)) return -1;
    int cc[ORC_MAXD], lu[ORC_MAXD], lv[ORC_MAXD];
    orc_cell K;
    cell_geometry(&D, cell, cc, &K, g_s1d, g_e1d);
    for (int iu = 0; iu < nloc; iu++) {
        int r = iu;
        for (int a = 0; a < d; a++) { lu[a] = r % (k + 1); r /= (k + 1); }
        for (int iv = 0; iv < nloc; iv++) {
            r = iv;
            for (int a = 0; a < d; a++) { lv[a] = r % (k + 1); r /= (k + 1); }
            double integral = 0;
            for (int i = 0; i < nint; i++)
                integral += orc_cell_integral_ex(&K, kinds[i], rule, lu, lv, wconst[i], NULL, NULL, NULL);
            out[(long)iv * nloc + iu] = integral;
        }
    }
    teardown_grid();
    return nloc;
}

/* RHS assembly: MatrixFESpace.writeCellIntegralsToRhs / addToVector
 * (M/basic/MatrixFESpace.java:44-57,124-129,141-145), cells ascending
 * (the reference's parallel loop is racy; sequential order is the defined one). */
int orc_assemble_rhs(int d, int k, const int *cells, const double *start, const double *end, int rule,
                     orc_sfun f, void *ctx, double *rhs)
{
    orc_asm_desc D;
    /* VALUE evaluates shapeFunction.value(x) (TPRightHandSideIntegral.java:26-29): in-cell only at
     * interior points, so Gauss-Lobatto rules are refused as in orc_assemble_w */
    if (!(rule >= 1 && rule <= 5)) return -1;
    memset(&D, 0, sizeof(D));
    D.d = d; D.k = k;
    long nc = 1;
    int nloc = 1;
    for (int a = 0; a < d; a++) { D.cells[a] = cells[a]; D.start[a] = start[a]; D.end[a] = end[a]; nc *= cells[a]; nloc *= k + 1; }
    if (setup_grid(&D)) return -1;
    int *cd = (int *)malloc(sizeof(int) * nc * nloc);
    const long n = orc_dof_table(d, k, cells, cd);
    for (long i = 0; i < n; i++) rhs[i] = 0;
    int lv[ORC_MAXD];
    for (long c = 0; c < nc; c++) {
        int cc[ORC_MAXD];
        orc_cell K;
        cell_geometry(&D, c, cc, &K, g_s1d, g_e1d);
        for (int iv = 0; iv < nloc; iv++) {
            int r = iv;
            for (int a = 0; a < d; a++) { lv[a] = r % (k + 1); r /= (k + 1); }
            const double integral = cell_rhs_integral(&K, rule, lv, f, ctx);
            if (integral != 0) rhs[cd[c * nloc + iv]] += integral;
        }
    }
    free(cd);
    teardown_grid();
    return 0;
}

/* Seeded sinusoid source f(x) = sum_m A_m prod_a sin(k_{a,m} pi x_a). */
typedef struct { int d, m; const int *kw; const double *amp; } orc_sines;
double orc_sines_eval(const double *x, int d, void *ctx)
{
    const orc_sines *S = (const orc_sines *)ctx;
    double s = 0;
    for (int j = 0; j < S->m; j++) {
        double t = S->amp[j];
        for (int a = 0; a < d; a++) t *= sin(S->kw[j * d + a] * M_PI * x[a]);
        s += t;
    }
    return s;
}
int orc_assemble_rhs_sines(int d, int k, const int *cells, const double *start, const double *end, int rule,
                           int m, const int *kw, const double *amp, double *rhs)
{
    orc_sines S = { d, m, kw, amp };
    return orc_assemble_rhs(d, k, cells, start, end, rule, orc_sines_eval, &S, rhs);
}

/* Node coordinates (LagrangeBasisFunction1D ctor :19, Cell1D.positionOnGrid :66-69):
 * computed in the cell where the node first appears. */
int orc_node_coords(int d, int k, const int *cells, const double *start, const double *end, double *xyz)
{
    orc_asm_desc D;
    memset(&D, 0, sizeof(D));
    D.d = d; D.k = k;
    long nc = 1;
    int nloc = 1;
    for (int a = 0; a < d; a++) { D.cells[a] = cells[a]; D.start[a] = start[a]; D.end[a] = end[a]; nc *= cells[a]; nloc *= k + 1; }
    if (setup_grid(&D)) return -1;
    int *cd = (int *)malloc(sizeof(int) * nc * nloc);
    const long n = orc_dof_table(d, k, cells, cd);
    char *done = (char *)calloc(n, 1);
    for (long c = 0; c < nc; c++) {
        int cc[ORC_MAXD];
        orc_cell K;
        cell_geometry(&D, c, cc, &K, g_s1d, g_e1d);
        for (int il = 0; il < nloc; il++) {
            const int g = cd[c * nloc + il];
            if (done[g]) continue;
            done[g] = 1;
            int r = il;
            for (int a = 0; a < d; a++) {
                const int l = r % (k + 1); r /= (k + 1);
                xyz[(long)g * d + a] = 1.0 / k * l * (K.e[a] - K.s[a]) + K.s[a];
            }
        }
    }
    free(done); free(cd);
    teardown_grid();
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Dirichlet elimination: AcceptsMatrixBoundaryValues.overWriteValue          */
/* (M/basic/AcceptsMatrixBoundaryValues.java:18-44) applied node by node in    */
/* the given order, on the CSR (entries kept as explicit zeros, diagonal = 1). */
/* Returns -1 if a node has no diagonal entry in the pattern.                  */
/* ------------------------------------------------------------------------- */
int orc_dirichlet(long n, const long *rowptr, const int *cols, double *vals, double *rhs,
                  long nb, const int *nodes, const double *g)
{
    /* column index: for each column, the positions of its entries */
    long nnz = rowptr[n];
    long *cptr = (long *)calloc(n + 1, sizeof(long));
    long *cpos = (long *)malloc(sizeof(long) * (nnz > 0 ? nnz : 1));
    int *crow = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    for (long p = 0; p < nnz; p++) cptr[cols[p] + 1]++;
    for (long c = 0; c < n; c++) cptr[c + 1] += cptr[c];
    long *fill = (long *)malloc(sizeof(long) * (n + 1));
    memcpy(fill, cptr, sizeof(long) * (n + 1));
    for (long r = 0; r < n; r++)
        for (long p = rowptr[r]; p < rowptr[r + 1]; p++) {
            const long q = fill[cols[p]]++;
            cpos[q] = p; crow[q] = (int)r;
        }
    char *fixed = (char *)calloc(n, 1);
    int rc = 0;
    for (long b = 0; b < nb; b++) {
        const int i = nodes[b];
        const double value = g ? g[b] : 0.0;
        fixed[i] = 1;
        for (long q = cptr[i]; q < cptr[i + 1]; q++) {          /* rhs -= column * value */
            const int j = crow[q];
            if (fixed[j]) continue;
            rhs[j] += -vals[cpos[q]] * value;
        }
        for (long q = cptr[i]; q < cptr[i + 1]; q++) vals[cpos[q]] = 0;   /* deleteColumn */
        long diag = -1;
        for (long p = rowptr[i]; p < rowptr[i + 1]; p++) {                 /* deleteRow */
            vals[p] = 0;
            if (cols[p] == i) diag = p;
        }
        if (diag < 0) { rc = -1; break; }
        vals[diag] = vals[diag] + (1 - vals[diag]);   /* set(1,i,i): appends 1 - at(i,i) */
        rhs[i] = value;
    }
    free(fixed); free(fill); free(crow); free(cpos); free(cptr);
    return rc;
}

/* Boundary nodes (getBoundaryNodes, AcceptsMatrixBoundaryValues.java:46-60):
 * the nodes lying on a boundary face, ascending.  Returns the count. */
long orc_boundary_nodes(int d, int k, const int *cells, int *out)
{
    long nc = 1;
    int nloc = 1;
    for (int a = 0; a < d; a++) { nc *= cells[a]; nloc *= k + 1; }
    int *cd = (int *)malloc(sizeof(int) * nc * nloc);
    const long n = orc_dof_table(d, k, cells, cd);
    char *bnd = (char *)calloc(n, 1);
    for (long c = 0; c < nc; c++) {
        int cc[ORC_MAXD];
        long t = c;
        for (int a = d - 1; a >= 0; a--) { cc[a] = (int)(t % cells[a]); t /= cells[a]; }
        for (int il = 0; il < nloc; il++) {
            int r = il, on = 0;
            for (int a = 0; a < d; a++) {
                const int p = cc[a] * k + r % (k + 1); r /= (k + 1);
                if (p == 0 || p == k * cells[a]) on = 1;
            }
            if (on) bnd[cd[c * nloc + il]] = 1;
        }
    }
    long nb = 0;
    for (long i = 0; i < n; i++) if (bnd[i]) out[nb++] = (int)i;
    free(bnd); free(cd);
    return nb;
}

/* ------------------------------------------------------------------------- */
/* Linear algebra                                                            */
/* ------------------------------------------------------------------------- */
/* Sparse product y = A x, row by row in column order (SparseMvMul.mvMul,
 * M/linalg/SparseMvMul.java:75-94). */
void orc_spmv(long n, const long *rowptr, const int *cols, const double *vals, const double *x, double *y)
{
    for (long r = 0; r < n; r++) {
        double acc = 0;
        for (long p = rowptr[r]; p < rowptr[r + 1]; p++) acc += vals[p] * x[cols[p]];
        y[r] = acc;
    }
}

/* y = A^T x (SparseMatrix.tvMul, M/linalg/SparseMatrix.java:491-501) */
void orc_spmv_t(long n, long ncols, const long *rowptr, const int *cols, const double *vals, const double *x, double *y)
{
    for (long c = 0; c < ncols; c++) y[c] = 0;
    for (long r = 0; r < n; r++)
        for (long p = rowptr[r]; p < rowptr[r + 1]; p++) y[cols[p]] += vals[p] * x[r];
}

typedef struct orc_op {
    long n;
    void (*apply)(struct orc_op *op, const double *x, double *y);
    void *ctx;
} orc_op;

typedef struct { long n; const long *rowptr; const int *cols; const double *vals; } csr_ctx;
static void csr_apply(orc_op *op, const double *x, double *y)
{
    const csr_ctx *c = (const csr_ctx *)op->ctx;
    orc_spmv(c->n, c->rowptr, c->cols, c->vals, x, y);
}
static void ident_apply(orc_op *op, const double *x, double *y)
{
    memcpy(y, x, sizeof(double) * op->n);
}

static double dot(long n, const double *a, const double *b)
{
    double r = 0;
    for (long i = 0; i < n; i++) r += a[i] * b[i];
    return r;
}
static double nrm(long n, const double *a) { return sqrt(dot(n, a, a)); }
static double *vnew(long n) { return (double *)calloc(n > 0 ? n : 1, sizeof(double)); }

/* Java 11 DoubleStream.sum(): Kahan-style compensated sum with the final
 * "sum + compensation" of java.util.stream.Collectors.computeFinalSum (JDK 11). */
typedef struct { double s, c, simple; } jsum;
static void jsum_add(jsum *j, double v)
{
    const double tmp = v - j->c;
    const double sum = j->s;
    const double velvel = sum + tmp;
    j->c = (velvel - sum) - tmp;
    j->s = velvel;
    j->simple += v;
}
static double jsum_final(const jsum *j)
{
    const double tmp = j->s + j->c;
    if (isnan(tmp) && isinf(j->simple)) return j->simple;
    return tmp;
}

typedef struct {
    int iterations;
    double resnorm;
    int fell_back;     /* BiCGStab -> GMRES fallback taken */
    int restarts;
} orc_result;

/* IterativeSolver.solveCG (M/linalg/IterativeSolver.java:48-100).
 * x in: start iterate (caller sets e_0 for the reference default); x out: solution. */
int orc_cg(orc_op *A, const double *rhs, double *x, double tol, long maxit, orc_result *res, double *hist)
{
    const long n = A->n;
    double *m = vnew(n), *r = vnew(n), *dfc = vnew(n), *z = vnew(n), *nr = vnew(n);
    A->apply(A, x, m);
    for (long i = 0; i < n; i++) r[i] = rhs[i] - m[i];
    memcpy(dfc, r, sizeof(double) * n);
    int it = 0;
    for (long iter = 0; iter < n && iter < maxit && nrm(n, r) > tol; iter++) {
        A->apply(A, dfc, z);
        const double alpha = dot(n, r, r) / dot(n, dfc, z);
        for (long i = 0; i < n; i++) x[i] += dfc[i] * alpha;
        for (long i = 0; i < n; i++) nr[i] = r[i] - z[i] * alpha;
        const double beta = dot(n, nr, nr) / dot(n, r, r);
        for (long i = 0; i < n; i++) dfc[i] = nr[i] + dfc[i] * beta;
        double *t = r; r = nr; nr = t;
        if (hist) hist[it] = nrm(n, r);
        it++;
    }
    res->iterations = it;
    res->resnorm = nrm(n, r);
    free(m); free(r); free(dfc); free(z); free(nr);
    return 0;
}

/* IterativeSolver.solvePCG (:102-162) */
int orc_pcg(orc_op *A, orc_op *M, const double *rhs, double *x, double tol, long maxit, orc_result *res, double *hist)
{
    const long n = A->n;
    double *t = vnew(n), *r = vnew(n), *z = vnew(n), *p = vnew(n), *Ap = vnew(n), *nr = vnew(n), *nz = vnew(n);
    A->apply(A, x, t);
    for (long i = 0; i < n; i++) r[i] = rhs[i] - t[i];
    M->apply(M, r, z);
    memcpy(p, z, sizeof(double) * n);
    int it = 0;
    for (long iter = 0; iter < n && iter < maxit && nrm(n, r) > tol; iter++) {
        A->apply(A, p, Ap);
        const double alpha = dot(n, r, z) / dot(n, p, Ap);
        for (long i = 0; i < n; i++) x[i] += p[i] * alpha;
        for (long i = 0; i < n; i++) nr[i] = r[i] - Ap[i] * alpha;
        M->apply(M, nr, nz);
        const double beta = dot(n, nr, nz) / dot(n, r, z);
        for (long i = 0; i < n; i++) p[i] = nz[i] + p[i] * beta;
        double *s = r; r = nr; nr = s;
        s = z; z = nz; nz = s;
        it++;
        if (hist) hist[it - 1] = nrm(n, r);
    }
    res->iterations = it;
    res->resnorm = nrm(n, r);
    free(t); free(r); free(z); free(p); free(Ap); free(nr); free(nz);
    return 0;
}

/* GMRES (M/linalg/GMRES.java:40-251).  H keeps the reference's COO "set"
 * semantics (a set appends value - at(i,j), SparseMatrix.java:259-271), which
 * a per-entry running sum reproduces exactly. */
typedef struct {
    int restart;       /* ITERATIONS_BEFORE_RESTART (150) */
    int max_restarts;  /* MAX_RESTARTS (200) */
    int restarts;
    int iterations;
} gm_state;

static void h_add(double *H, int ld, int i, int j, double v) { if (v != 0) H[i * ld + j] = H[i * ld + j] + v; }
static void h_set(double *H, int ld, int i, int j, double v) { H[i * ld + j] = H[i * ld + j] + (v - H[i * ld + j]); }

static void gm_alpha(const double *gamma, const double *H, int ld, int j, double *alpha)
{
    for (int i = j; i >= 0; i--) {
        jsum s = { 0, 0, 0 };
        for (int k = i + 1; k <= j; k++) jsum_add(&s, H[i * ld + k] * alpha[k]);
        alpha[i] = 1. / H[i * ld + i] * (gamma[i] - jsum_final(&s));
    }
}

/* one GMRES cycle; returns 1 if a restart is needed (and x updated). M may be NULL. */
static int gm_cycle(gm_state *st, orc_op *M, orc_op *A, const double *b, double *x, double tol)
{
    const long n = A->n;
    const int ld = st->restart + 10;
    double *t = vnew(n), *r0 = vnew(n), *r = vnew(n);
    A->apply(A, x, t);
    for (long i = 0; i < n; i++) r0[i] = b[i] - t[i];
    if (M) M->apply(M, r0, r); else memcpy(r, r0, sizeof(double) * n);
    if (nrm(n, r) <= tol) { free(t); free(r0); free(r); return 0; }
    double *gamma = vnew(ld), *H = vnew((long)ld * ld), *c = vnew(n > ld ? n : ld), *s = vnew(n > ld ? n : ld);
    const int maxv = ld + 1;
    double **V = (double **)calloc(maxv, sizeof(double *));
    int nv = 0;
    gamma[0] = nrm(n, r);
    V[nv] = vnew(n);
    for (long i = 0; i < n; i++) V[nv][i] = r[i] * (1. / gamma[0]);
    nv++;
    double *q = vnew(n), *tq = vnew(n), *w = vnew(n), *f = vnew(maxv), *orth = vnew(n);
    int j;
    for (j = 0; j < n; j++) {                                    /* innerGMRES :136-176 */
        if (M) { A->apply(A, V[j], tq); M->apply(M, tq, q); } else A->apply(A, V[j], q);
        for (int i = 0; i < nv; i++) f[i] = dot(n, V[i], q);
        for (long i = 0; i < n; i++) orth[i] = 0;
        for (int i = 0; i < nv; i++) for (long e = 0; e < n; e++) orth[e] = orth[e] + V[i][e] * f[i];
        for (long e = 0; e < n; e++) w[e] = q[e] - orth[e];
        for (int i = 0; i < nv; i++) h_add(H, ld, i, j, f[i]);   /* addColumn */
        h_add(H, ld, j + 1, j, nrm(n, w));
        for (int i = 0; i < j; i++) {                            /* calculateGivens */
            const double ci = c[i], si = s[i];
            const double hij = H[i * ld + j], hipj = H[(i + 1) * ld + j];
            h_set(H, ld, i, j, hij * ci + hipj * si);
            h_set(H, ld, i + 1, j, -hij * si + hipj * ci);
        }
        const double hjj = H[j * ld + j], hj1 = H[(j + 1) * ld + j];
        const double beta = sqrt(pow(hjj, 2) + pow(hj1, 2));
        s[j] = H[(j + 1) * ld + j] / beta;
        c[j] = H[j * ld + j] / beta;
        h_set(H, ld, j, j, beta);
        gamma[j + 1] = -s[j] * gamma[j];
        gamma[j] = c[j] * gamma[j];
        if (fabs(gamma[j + 1]) < tol || j > st->restart) break;
        V[nv] = vnew(n);
        const double hn = H[(j + 1) * ld + j];
        for (long e = 0; e < n; e++) V[nv][e] = w[e] * (1. / hn);
        nv++;
        st->iterations++;
        if (nv >= maxv) { j++; break; }
    }
    if (j == n) j--;
    double *alpha = vnew(j + 1);
    gm_alpha(gamma, H, ld, j, alpha);
    double *av = vnew(n);
    for (int i = 0; i <= j; i++) for (long e = 0; e < n; e++) av[e] = av[e] + V[i][e] * alpha[i];
    double *res = vnew(n);
    for (long e = 0; e < n; e++) res[e] = x[e] + av[e];
    A->apply(A, res, t);
    for (long e = 0; e < n; e++) t[e] = t[e] - b[e];
    const double residual = nrm(n, t);
    int again = 0;
    if ((j > st->restart && st->restarts < st->max_restarts) || residual > tol) {
        st->restarts++;
        again = 1;
    }
    for (long e = 0; e < n; e++) x[e] = x[e] + av[e];
    for (int i = 0; i < nv; i++) free(V[i]);
    free(V); free(alpha); free(av); free(res); free(q); free(tq); free(w); free(f); free(orth);
    free(gamma); free(H); free(c); free(s); free(t); free(r0); free(r);
    return again;
}

/* GMRES.solve with optional left preconditioner M; x in: start iterate.
 * max_cycles bounds the (otherwise unbounded) restart recursion. */
int orc_gmres(orc_op *A, orc_op *M, const double *b, double *x, double tol, int restart, int max_cycles, orc_result *out)
{
    gm_state st = { restart, 200, 0, 0 };
    int cycles = 0;
    while (gm_cycle(&st, M, A, b, x, tol) && ++cycles < max_cycles) {}
    out->iterations = st.iterations;
    out->restarts = st.restarts;
    const long n = A->n;
    double *t = vnew(n);
    A->apply(A, x, t);
    for (long e = 0; e < n; e++) t[e] = b[e] - t[e];
    out->resnorm = nrm(n, t);
    free(t);
    return 0;
}

/* IterativeSolver.solveBiCGStab / solvePBiCGStab (:256-402).  M NULL = unpreconditioned.
 * On the rho breakdown the reference calls (P)GMRES on the original rhs. */
int orc_bicgstab(orc_op *A, orc_op *M, const double *rhs, double *x, double tol, orc_result *res)
{
    const long n = A->n;
    double *v = vnew(n), *s = vnew(n), *t = vnew(n), *r0 = vnew(n), *r = vnew(n), *p = vnew(n);
    double *y = vnew(n), *z = vnew(n), *pt = vnew(n), *tmp = vnew(n);
    A->apply(A, x, tmp);
    for (long i = 0; i < n; i++) r0[i] = rhs[i] - tmp[i];
    memcpy(r, r0, sizeof(double) * n);
    double alpha = 1, omega = 1, rho = 1, rhoLast = 1, beta = 1;
    int it = 0;
    res->fell_back = 0;
    const long maxit = M ? n * 10 : n;
    for (long iter = 0; iter < maxit && nrm(n, r) > tol; iter++) {
        rho = dot(n, r, r0);
        if (fabs(rhoLast) < tol) {
            res->fell_back = 1;
            if (M) { for (long i = 0; i < n; i++) x[i] = 0; }
            else { for (long i = 0; i < n; i++) x[i] = 0; x[0] = 1; }
            orc_result g;
            orc_gmres(A, M, rhs, x, tol, 150, 1000, &g);
            res->iterations = g.iterations;
            res->resnorm = g.resnorm;
            goto done;
        }
        beta = alpha / omega * rho / rhoLast;
        for (long i = 0; i < n; i++) p[i] = (r[i] + p[i] * beta) - v[i] * (omega * beta);
        if (M) { M->apply(M, p, y); A->apply(A, y, v); } else A->apply(A, p, v);
        alpha = rho / dot(n, v, r0);
        for (long i = 0; i < n; i++) s[i] = r[i] - v[i] * alpha;
        if (nrm(n, s) < tol) {
            const double *dir = M ? y : p;
            for (long i = 0; i < n; i++) x[i] = x[i] + dir[i] * alpha;
            memcpy(r, s, sizeof(double) * n);
            break;
        }
        if (M) {
            M->apply(M, s, z); A->apply(A, z, t); M->apply(M, t, pt);
            omega = dot(n, pt, z) / dot(n, pt, pt);
            for (long i = 0; i < n; i++) x[i] = (x[i] + y[i] * alpha) + z[i] * omega;
        } else {
            A->apply(A, s, t);
            omega = dot(n, t, s) / dot(n, t, t);
            for (long i = 0; i < n; i++) x[i] = (x[i] + p[i] * alpha) + s[i] * omega;
        }
        for (long i = 0; i < n; i++) r[i] = s[i] - t[i] * omega;
        rhoLast = rho;
        it++;
    }
    res->iterations = it;
    A->apply(A, x, tmp);
    for (long i = 0; i < n; i++) tmp[i] = rhs[i] - tmp[i];
    res->resnorm = nrm(n, tmp);
done:
    free(v); free(s); free(t); free(r0); free(r); free(p); free(y); free(z); free(pt); free(tmp);
    return 0;
}

/* Dense LU with partial pivoting (textbook, Golub & Van Loan alg. 3.4.1),
 * the coarse solve that the reference delegates to UJMP (DenseMatrix.solve,
 * M/linalg/DenseMatrix.java:491-519).  A is n x n row-major, overwritten. */
int orc_lu_solve(int n, double *A, const double *b, double *x)
{
    int *piv = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) piv[i] = i;
    for (int kk = 0; kk < n; kk++) {
        int pr = kk;
        double best = fabs(A[(long)kk * n + kk]);
        for (int i = kk + 1; i < n; i++) if (fabs(A[(long)i * n + kk]) > best) { best = fabs(A[(long)i * n + kk]); pr = i; }
        if (best == 0) { free(piv); return -1; }
        if (pr != kk) {
            for (int j = 0; j < n; j++) { double t = A[(long)kk * n + j]; A[(long)kk * n + j] = A[(long)pr * n + j]; A[(long)pr * n + j] = t; }
            int t = piv[kk]; piv[kk] = piv[pr]; piv[pr] = t;
        }
        for (int i = kk + 1; i < n; i++) {
            const double l = A[(long)i * n + kk] / A[(long)kk * n + kk];
            A[(long)i * n + kk] = l;
            for (int j = kk + 1; j < n; j++) A[(long)i * n + j] -= l * A[(long)kk * n + j];
        }
    }
    for (int i = 0; i < n; i++) {
        double s = b[piv[i]];
        for (int j = 0; j < i; j++) s -= A[(long)i * n + j] * x[j];
        x[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = x[i];
        for (int j = i + 1; j < n; j++) s -= A[(long)i * n + j] * x[j];
        x[i] = s / A[(long)i * n + i];
    }
    free(piv);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Geometric multigrid: MGPreconditionerInterface (M/multigrid/               */
/* MGPreconditionerInterface.java:125-198 mgStepV, :293-326 vCycle and        */
/* fullVCycleCorrection, :352-379 mvMul), MGPreconditionerSpace prolongation   */
/* (M/multigrid/MGPreconditionerSpace.java:59-109), smoothers Jacobi          */
/* (JacobiSmoother.java:12-50), Richardson (RichardsonSmoother.java:13-50) and */
/* Chebyshev-Jacobi (not in the reference; Saad, "Iterative Methods for Sparse */
/* Linear Systems", 2nd ed., Alg. 12.1, applied to D^-1 A).                   */
/* ------------------------------------------------------------------------- */
enum { ORC_SM_RICHARDSON = 0, ORC_SM_JACOBI = 1, ORC_SM_CHEBYSHEV = 2 };

typedef struct {
    int kind;
    int runs;          /* Richardson/Jacobi runs; Chebyshev: number of polynomial applications */
    double omega;      /* Richardson/Jacobi relaxation */
    int degree;        /* Chebyshev degree */
    double lmin, lmax; /* Chebyshev interval on D^-1 A */
} orc_smoother;

typedef struct {
    long n;
    const long *rowptr; const int *cols; const double *vals;   /* A_l after Dirichlet */
    const unsigned char *bnd;                                  /* boundary mask */
    double *dinv;                                              /* Jacobi diagonal inverse */
    /* prolongation from this level to level+1 (fine rows, coarse cols) */
    long pn; const long *prow; const int *pcol; const double *pval;
    orc_smoother sm;
} orc_level;

typedef struct {
    int nlev;
    orc_level *lev;     /* 0 = coarsest */
    double *coarse_dense;  /* dense copy of A_0 */
    int cycles;         /* V-cycles after the FMG correction (reference: 3) */
    int fmg;            /* 1 = reference mvMul (copy BC, FMG, cycles V); 0 = cycles V from zero;
                           2 = MGPReconditioner2.mvMul: v + FMG solve of v - A v */
    int raw;            /* 1 = MGSpace transfers: the prolongation matrix without zero boundary values */
} orc_mg;

static void spmv_l(const orc_level *L, const double *x, double *y) { orc_spmv(L->n, L->rowptr, L->cols, L->vals, x, y); }

/* Test hook: plants a known mistake in smooth() so that the closed-form smoother tests can show
 * they catch it.  0 = none (the default, never set outside those tests); 1 = Chebyshev c1 without
 * the factor rho_old; 2 = Chebyshev c2 without the factor 2; 3 = interval centre taken as the
 * half-width; 4 = Jacobi without omega; 5 = Richardson without its residual-norm exit. */
static int g_smooth_plant = 0;
void orc_set_smooth_plant(int plant) { g_smooth_plant = plant; }

static void smooth(const orc_level *L, const double *b, double *x)
{
    const long n = L->n;
    double *t = vnew(n), *r = vnew(n);
    const orc_smoother *S = &L->sm;
    if (S->kind == ORC_SM_RICHARDSON) {
        for (int it = 0; it < S->runs; it++) {
            spmv_l(L, x, t);
            for (long i = 0; i < n; i++) r[i] = b[i] - t[i];
            if (nrm(n, r) < 1e-10 && g_smooth_plant != 5) break;
            for (long i = 0; i < n; i++) x[i] = x[i] + r[i] * S->omega;
        }
    } else if (S->kind == ORC_SM_JACOBI) {
        for (int it = 0; it < S->runs; it++) {
            spmv_l(L, x, t);
            for (long i = 0; i < n; i++) r[i] = b[i] - t[i];
            const double om = g_smooth_plant == 4 ? 1.0 : S->omega;
            for (long i = 0; i < n; i++) x[i] = x[i] + (L->dinv[i] * r[i]) * om;
        }
    } else {
        const double delta = 0.5 * (S->lmax - S->lmin);
        const double theta = g_smooth_plant == 3 ? delta : 0.5 * (S->lmax + S->lmin);
        const double sigma = theta / delta;
        double *dd = vnew(n);
        for (int app = 0; app < S->runs; app++) {
            double rho_old = 1.0 / sigma;
            for (int kdeg = 0; kdeg < S->degree; kdeg++) {
                spmv_l(L, x, t);
                for (long i = 0; i < n; i++) r[i] = b[i] - t[i];
                if (kdeg == 0) {
                    for (long i = 0; i < n; i++) dd[i] = (L->dinv[i] * r[i]) / theta;
                } else {
                    const double rho = 1.0 / (2.0 * sigma - rho_old);
                    const double c1 = g_smooth_plant == 1 ? rho : rho * rho_old;
                    const double c2 = (g_smooth_plant == 2 ? 1.0 : 2.0) * rho / delta;
                    for (long i = 0; i < n; i++) dd[i] = c1 * dd[i] + c2 * (L->dinv[i] * r[i]);
                    rho_old = rho;
                }
                for (long i = 0; i < n; i++) x[i] = x[i] + dd[i];
            }
        }
        free(dd);
    }
    free(t); free(r);
}

static void prolong(const orc_mg *G, int l, const double *xc, double *xf, int zero_bc)
{
    const orc_level *C = &G->lev[l], *F = &G->lev[l + 1];
    orc_spmv(C->pn, C->prow, C->pcol, C->pval, xc, xf);
    if (zero_bc && !G->raw) for (long i = 0; i < F->n; i++) if (F->bnd[i]) xf[i] = 0;
}
static void restrict_(const orc_mg *G, int l, const double *xf, double *xc)
{
    const orc_level *C = &G->lev[l];
    orc_spmv_t(C->pn, C->n, C->prow, C->pcol, C->pval, xf, xc);
    if (!G->raw) for (long i = 0; i < C->n; i++) if (C->bnd[i]) xc[i] = 0;
}

static void coarse_solve(const orc_mg *G, const double *b, double *x)
{
    const long n = G->lev[0].n;
    double *A = vnew(n * n);
    memcpy(A, G->coarse_dense, sizeof(double) * n * n);
    orc_lu_solve((int)n, A, b, x);
    free(A);
}

/* mgStepV(level, guess, rhs); guess is updated in place */
static void mg_step_v(const orc_mg *G, int l, double *x, const double *b)
{
    const orc_level *L = &G->lev[l];
    const long n = L->n;
    if (l == 0) { coarse_solve(G, b, x); return; }
    smooth(L, b, x);
    double *t = vnew(n), *rc = vnew(G->lev[l - 1].n), *ec = vnew(G->lev[l - 1].n), *pe = vnew(n);
    spmv_l(L, x, t);
    for (long i = 0; i < n; i++) t[i] = b[i] - t[i];
    restrict_(G, l - 1, t, rc);
    mg_step_v(G, l - 1, ec, rc);
    prolong(G, l - 1, ec, pe, 1);
    for (long i = 0; i < n; i++) x[i] = x[i] + pe[i];
    smooth(L, b, x);
    free(t); free(rc); free(ec); free(pe);
}

void orc_mg_vcycle(const orc_mg *G, double *x, const double *b) { mg_step_v(G, G->nlev - 1, x, b); }

/* MGSpace (M/multigrid/MGSpace.java:25-206): transfers by the raw prolongation matrix */
void orc_mg_set_raw(orc_mg *G, int raw) { G->raw = raw; }

/* mgStepW (MGPreconditionerInterface.java:200-291): pre-smooth, then twice
 * { restrict the defect, V-cycle (mgStepV) on the next coarser level from zero,
 *   prolong and add }, post-smooth; wCycle (:300-305) starts it on the finest level */
static void mg_step_w(const orc_mg *G, int l, double *x, const double *b)
{
    const orc_level *L = &G->lev[l];
    const long n = L->n;
    if (l == 0) { coarse_solve(G, b, x); return; }
    smooth(L, b, x);
    double *t = vnew(n), *rc = vnew(G->lev[l - 1].n), *ec = vnew(G->lev[l - 1].n), *pe = vnew(n);
    for (int k = 0; k < 2; k++) {
        spmv_l(L, x, t);
        for (long i = 0; i < n; i++) t[i] = b[i] - t[i];
        restrict_(G, l - 1, t, rc);
        for (long i = 0; i < G->lev[l - 1].n; i++) ec[i] = 0;
        mg_step_v(G, l - 1, ec, rc);
        prolong(G, l - 1, ec, pe, 1);
        for (long i = 0; i < n; i++) x[i] = x[i] + pe[i];
    }
    smooth(L, b, x);
    free(t); free(rc); free(ec); free(pe);
}

void orc_mg_wcycle(const orc_mg *G, double *x, const double *b) { mg_step_w(G, G->nlev - 1, x, b); }

/* MGPreconditionerInterface.mvMul (:352-379), or plain cycles of V from zero when fmg == 0 */
void orc_mg_apply(const orc_mg *G, const double *v, double *out)
{
    const int L = G->nlev - 1;
    const orc_level *F = &G->lev[L];
    const long n = F->n;
    if (!G->fmg) {
        for (long i = 0; i < n; i++) out[i] = 0;
        for (int c = 0; c < G->cycles; c++) mg_step_v(G, L, out, v);
        return;
    }
    double *init = vnew(n), *t = vnew(n), *defect = vnew(n);
    if (G->fmg == 2)   /* MGPReconditioner2.mvMul (M/multigrid/MGPReconditioner2.java:38-76): v + fullVCycleSolver(v - A v) */
        memcpy(init, v, sizeof(double) * n);
    else
        for (long i = 0; i < n; i++) if (F->bnd[i]) init[i] = v[i];      /* copyBoundaryValues */
    spmv_l(F, init, t);
    for (long i = 0; i < n; i++) defect[i] = v[i] - t[i];
    /* fullVCycleCorrection (:307-326) */
    double **rhs = (double **)calloc(G->nlev, sizeof(double *));
    rhs[L] = vnew(n);
    memcpy(rhs[L], defect, sizeof(double) * n);
    for (int l = L - 1; l >= 0; l--) { rhs[l] = vnew(G->lev[l].n); restrict_(G, l, rhs[l + 1], rhs[l]); }
    double *it = vnew(G->lev[0].n);
    for (int l = 0; l < L; l++) {
        mg_step_v(G, l, it, rhs[l]);
        double *nx = vnew(G->lev[l + 1].n);
        prolong(G, l, it, nx, 0);                       /* raw prolongation matrix */
        free(it);
        it = nx;
    }
    for (int c = 0; c < G->cycles; c++) mg_step_v(G, L, it, defect);
    for (long i = 0; i < n; i++) out[i] = init[i] + it[i];
    for (int l = 0; l <= L; l++) free(rhs[l]);
    free(rhs); free(it); free(init); free(t); free(defect);
}

static void mg_apply_op(orc_op *op, const double *x, double *y) { orc_mg_apply((const orc_mg *)op->ctx, x, y); }

/* Prolongation matrix between two nested Qk spaces of the same box
 * (MGPreconditionerSpace.createProlongationMatrix :59-76 with
 * getRefinedFunctions, MGPreconditionerInterface.java:60-86): P[g,f] = value of
 * coarse function f at the node of fine function g, entries equal to 0.0
 * skipped.  The coarse function is evaluated in the first coarse cell (lex
 * order) of its support that contains the node.  Returns a CSR (fine x coarse). */
orc_csr *orc_prolongation(int d, int k, const int *cells_c, const double *start, const double *end)
{
    int cells_f[ORC_MAXD] = {0, 0, 0};
    long ncc = 1;
    int nloc = 1;
    for (int a = 0; a < d; a++) { cells_f[a] = 2 * cells_c[a]; ncc *= cells_c[a]; nloc *= k + 1; }
    long nf = 1;
    for (int a = 0; a < d; a++) nf *= (long)k * cells_f[a] + 1;
    double *xf = vnew(nf * d);
    orc_node_coords(d, k, cells_f, start, end, xf);
    int *cdc = (int *)malloc(sizeof(int) * ncc * nloc);
    orc_dof_table(d, k, cells_c, cdc);
    /* coarse cell geometry */
    double sc[ORC_MAXD][4096], ec[ORC_MAXD][4096];
    for (int a = 0; a < d; a++) orc_cells1d(start[a], end[a], cells_c[a], sc[a], ec[a]);
    /* fine DOF -> lattice (via fine dof table) */
    long ncf = 1;
    for (int a = 0; a < d; a++) ncf *= cells_f[a];
    int *cdf = (int *)malloc(sizeof(int) * ncf * nloc);
    orc_dof_table(d, k, cells_f, cdf);
    int *flat = (int *)malloc(sizeof(int) * nf * d);
    for (long c = 0; c < ncf; c++) {
        int cc[ORC_MAXD];
        long t = c;
        for (int a = d - 1; a >= 0; a--) { cc[a] = (int)(t % cells_f[a]); t /= cells_f[a]; }
        for (int il = 0; il < nloc; il++) {
            int r = il;
            const int g = cdf[c * nloc + il];
            for (int a = 0; a < d; a++) { flat[(long)g * d + a] = cc[a] * k + r % (k + 1); r /= (k + 1); }
        }
    }
    coo_t *coo = (coo_t *)malloc(sizeof(coo_t) * (nf * (long)nloc * (1 << d) + 1));
    long m = 0;
    coo[m].row = 0; coo[m].col = 0; coo[m].val = 0.0; coo[m].seq = 0; m++;
    /* for each fine node: coarse cells containing its position, lowest first; the
     * coarse functions with support there are the nodes of that coarse cell. */
    for (long g = 0; g < nf; g++) {
        const double *x = &xf[g * d];
        /* candidate coarse functions: nodes of the coarse cells (lex order) whose
         * closure contains x; each coarse function is evaluated once, in the first
         * such cell. */
        int lo[ORC_MAXD], hi[ORC_MAXD];
        for (int a = 0; a < d; a++) {
            const int fl = flat[g * d + a];
            /* coarse cells touching fine lattice fl: fine cell index ~ fl/k, coarse = fine/2 */
            int c0 = (fl / k) / 2, c1 = c0;
            if (fl % (2 * k) == 0) { c0 = fl / (2 * k) - 1; c1 = fl / (2 * k); }
            if (c0 < 0) c0 = 0;
            if (c1 >= cells_c[a]) c1 = cells_c[a] - 1;
            if (c0 >= cells_c[a]) c0 = cells_c[a] - 1;
            lo[a] = c0; hi[a] = c1;
        }
        long done_cnt = 0;
        int done_f[1024];
        int cnt[ORC_MAXD];
        for (int a = 0; a < d; a++) cnt[a] = hi[a] - lo[a] + 1;
        int ncand = 1;
        for (int a = 0; a < d; a++) ncand *= cnt[a];
        for (int ci = 0; ci < ncand; ci++) {
            int cc[ORC_MAXD], t = ci;
            for (int a = d - 1; a >= 0; a--) { cc[a] = lo[a] + t % cnt[a]; t /= cnt[a]; }
            long clin = 0;
            for (int a = 0; a < d; a++) clin = clin * cells_c[a] + cc[a];
            int in = 1;
            for (int a = 0; a < d; a++) if (!cell1d_in(sc[a][cc[a]], ec[a][cc[a]], x[a])) in = 0;
            if (!in) continue;
            for (int il = 0; il < nloc; il++) {
                const int f = cdc[clin * nloc + il];
                int seen = 0;
                for (long q = 0; q < done_cnt; q++) if (done_f[q] == f) seen = 1;
                if (seen) continue;
                done_f[done_cnt++] = f;
                int r = il;
                double val = 1;
                for (int a = 0; a < d; a++) {
                    const int l = r % (k + 1); r /= (k + 1);
                    val *= lag_val_cell(sc[a][cc[a]], ec[a][cc[a]], l, k, x[a]);
                }
                if (val != 0) { coo[m].row = (int)g; coo[m].col = f; coo[m].val = val; coo[m].seq = m; m++; }
            }
        }
    }
    orc_csr *P = coo_to_csr(nf, coo, m);
    free(coo); free(flat); free(cdf); free(cdc); free(xf);
    return P;
}

/* MG hierarchy over a box: levels with base_cells * 2^l cells, l = 0..nlev-1,
 * each assembled with the same integrals (re-discretisation as in the
 * reference's createSystem) and zero Dirichlet values on the whole boundary. */
orc_mg *orc_mg_create(int d, int k, const int *base_cells, int nlev, const double *start, const double *end,
                      int nint, const int *kinds, const double *wconst, int rule,
                      int smoother_kind, int runs, double omega, int degree, const double *lmin, const double *lmax,
                      int cycles, int fmg)
{
    orc_mg *G = (orc_mg *)calloc(1, sizeof(orc_mg));
    G->nlev = nlev; G->cycles = cycles; G->fmg = fmg;
    G->lev = (orc_level *)calloc(nlev, sizeof(orc_level));
    for (int l = 0; l < nlev; l++) {
        int cells[ORC_MAXD];
        for (int a = 0; a < d; a++) cells[a] = base_cells[a] << l;
        orc_csr *A = orc_assemble(d, k, cells, start, end, nint, kinds, wconst, NULL, rule);
        orc_level *L = &G->lev[l];
        L->n = A->n;
        int *bn = (int *)malloc(sizeof(int) * A->n);
        const long nb = orc_boundary_nodes(d, k, cells, bn);
        double *rhs = vnew(A->n);
        orc_dirichlet(A->n, A->rowptr, A->cols, A->vals, rhs, nb, bn, NULL);
        unsigned char *mask = (unsigned char *)calloc(A->n, 1);
        for (long i = 0; i < nb; i++) mask[bn[i]] = 1;
        L->bnd = mask;
        L->rowptr = A->rowptr; L->cols = A->cols; L->vals = A->vals;
        L->dinv = vnew(A->n);
        for (long i = 0; i < A->n; i++) {                     /* JacobiSmoother ctor :14-25 */
            double dg = 0;
            for (long p = A->rowptr[i]; p < A->rowptr[i + 1]; p++) if (A->cols[p] == i) dg += A->vals[p];
            if (fabs(dg) < 1e-10) dg = 1;
            L->dinv[i] = 1. / dg;
        }
        L->sm.kind = smoother_kind; L->sm.runs = runs; L->sm.omega = omega; L->sm.degree = degree;
        L->sm.lmin = lmin ? lmin[l] : 0; L->sm.lmax = lmax ? lmax[l] : 0;
        free(A);   /* arrays are now owned by the level */
        free(bn); free(rhs);
        if (l > 0) {
            int cc[ORC_MAXD];
            for (int a = 0; a < d; a++) cc[a] = base_cells[a] << (l - 1);
            orc_csr *P = orc_prolongation(d, k, cc, start, end);
            orc_level *C = &G->lev[l - 1];
            C->pn = P->n; C->prow = P->rowptr; C->pcol = P->cols; C->pval = P->vals;
            free(P);
        }
    }
    const long n0 = G->lev[0].n;
    G->coarse_dense = vnew(n0 * n0);
    for (long i = 0; i < n0; i++)
        for (long p = G->lev[0].rowptr[i]; p < G->lev[0].rowptr[i + 1]; p++)
            G->coarse_dense[i * n0 + G->lev[0].cols[p]] += G->lev[0].vals[p];
    return G;
}

void orc_mg_free(orc_mg *G)
{
    if (!G) return;
    for (int l = 0; l < G->nlev; l++) {
        orc_level *L = &G->lev[l];
        free((void *)L->rowptr); free((void *)L->cols); free((void *)L->vals);
        free((void *)L->bnd); free(L->dinv);
        free((void *)L->prow); free((void *)L->pcol); free((void *)L->pval);
    }
    free(G->lev); free(G->coarse_dense); free(G);
}

/* Accessors for tests */
long orc_mg_level_n(const orc_mg *G, int l) { return G->lev[l].n; }
void orc_mg_level_dinv(const orc_mg *G, int l, double *out) { memcpy(out, G->lev[l].dinv, sizeof(double) * G->lev[l].n); }
long orc_mg_level_nnz(const orc_mg *G, int l) { return G->lev[l].rowptr[G->lev[l].n]; }
void orc_mg_level_csr(const orc_mg *G, int l, long *rowptr, int *cols, double *vals)
{
    const orc_level *L = &G->lev[l];
    memcpy(rowptr, L->rowptr, sizeof(long) * (L->n + 1));
    memcpy(cols, L->cols, sizeof(int) * L->rowptr[L->n]);
    memcpy(vals, L->vals, sizeof(double) * L->rowptr[L->n]);
}
long orc_mg_prol_nnz(const orc_mg *G, int l) { return G->lev[l].prow[G->lev[l].pn]; }
void orc_mg_prol_csr(const orc_mg *G, int l, long *rowptr, int *cols, double *vals)
{
    const orc_level *L = &G->lev[l];
    memcpy(rowptr, L->prow, sizeof(long) * (L->pn + 1));
    memcpy(cols, L->pcol, sizeof(int) * L->prow[L->pn]);
    memcpy(vals, L->pval, sizeof(double) * L->prow[L->pn]);
}
void orc_mg_smooth(const orc_mg *G, int l, const double *b, double *x) { smooth(&G->lev[l], b, x); }
void orc_mg_set_interval(orc_mg *G, int l, double lmin, double lmax) { G->lev[l].sm.lmin = lmin; G->lev[l].sm.lmax = lmax; }

/* ------------------------------------------------------------------------- */
/* Solver entry points over CSR / dense / MG operators, for ctypes.           */
/* precond: 0 = none, 1 = identity, 2 = MG (mg != NULL), 3 = CSR matrix M     */
/* ------------------------------------------------------------------------- */
typedef struct {
    int iterations, fell_back, restarts;
    double resnorm;
} orc_solve_out;

/* GMRES2.solveInternal (M/linalg/GMRES2.java:59-104) with Arnoldi (M/linalg/Arnoldi.java:13-82),
 * M = identity when NULL (GMRES2.solve(A, b) uses SparseMatrix.identity).  MA = M(A .):
 *   Arnoldi(Mr): q0 = Mr, w0 = q0/|q0|, q1 = MA w0, v0 = w0 * (1/|q1|), Av0 = q1 * (1/|q1|),
 *                q2 = MA (q1/|q1|)
 *   while |r| > tol:  t = (Mr.Av)/(Av.Av) for the latest (v, Av); x += v t; r -= (A v) t;
 *                     Mr -= Av t; arnoldiStep: w = q/|q| (latest q), q' = MA w, Aw = q';
 *                     for every (v_i, Av_i): p = Aw.Av_i, w -= v_i p, Aw -= Av_i p;
 *                     append w * (1/|Aw|), Aw * (1/|Aw|); restart (from the new iterate,
 *                     recomputing r and Mr) when |Aw| < 1e-13.
 * The stop test uses the recursively updated residual, absolute tol.  maxit bounds the
 * iterations, which the reference leaves unbounded. */
static void ma_apply(orc_op *A, orc_op *M, const double *x, double *tmp, double *y)
{
    A->apply(A, x, tmp);
    if (M) M->apply(M, tmp, y); else memcpy(y, tmp, sizeof(double) * A->n);
}

int orc_gmres2(orc_op *A, orc_op *M, const double *b, double *x, double tol, long maxit, orc_result *out)
{
    const long n = A->n;
    double *r = vnew(n), *Mr = vnew(n), *t = vnew(n), *q = vnew(n), *qn = vnew(n), *w = vnew(n), *aw = vnew(n);
    long cap = 16;
    double **V = (double **)calloc(cap, sizeof(double *)), **AV = (double **)calloc(cap, sizeof(double *));
    long it = 0;
    int restarts = 0;
    for (;;) {
        A->apply(A, x, t);
        for (long e = 0; e < n; e++) r[e] = b[e] - t[e];
        if (M) M->apply(M, r, Mr); else memcpy(Mr, r, sizeof(double) * n);
        /* Arnoldi constructor */
        long nv = 0;
        double s = 1. / nrm(n, Mr);
        for (long e = 0; e < n; e++) w[e] = Mr[e] * s;                  /* w0 */
        ma_apply(A, M, w, t, q);                                          /* q1 = Aw0 */
        const double n1 = nrm(n, q);
        V[0] = vnew(n); AV[0] = vnew(n);
        for (long e = 0; e < n; e++) { V[0][e] = w[e] * (1. / n1); AV[0][e] = q[e] * (1. / n1); }
        nv = 1;
        s = 1. / n1;
        for (long e = 0; e < n; e++) w[e] = q[e] * s;
        ma_apply(A, M, w, t, qn);                                         /* q2 */
        memcpy(q, qn, sizeof(double) * n);
        int restart = 0;
        while (nrm(n, r) > tol) {
            if (it >= maxit) break;
            const double *v = V[nv - 1], *av = AV[nv - 1];
            const double tt = dot(n, Mr, av) / dot(n, av, av);
            for (long e = 0; e < n; e++) x[e] = x[e] + v[e] * tt;
            A->apply(A, v, t);
            for (long e = 0; e < n; e++) r[e] = r[e] - t[e] * tt;
            for (long e = 0; e < n; e++) Mr[e] = Mr[e] - av[e] * tt;
            it++;
            /* arnoldiStep */
            s = 1. / nrm(n, q);
            for (long e = 0; e < n; e++) w[e] = q[e] * s;
            ma_apply(A, M, w, t, qn);
            memcpy(aw, qn, sizeof(double) * n);
            for (long i = 0; i < nv; i++) {
                const double p = dot(n, aw, AV[i]);
                for (long e = 0; e < n; e++) w[e] = w[e] - V[i][e] * p;
                for (long e = 0; e < n; e++) aw[e] = aw[e] - AV[i][e] * p;
            }
            const double na = nrm(n, aw);
            if (nv == cap) {
                cap *= 2;
                V = (double **)realloc(V, sizeof(double *) * cap);
                AV = (double **)realloc(AV, sizeof(double *) * cap);
            }
            V[nv] = vnew(n); AV[nv] = vnew(n);
            for (long e = 0; e < n; e++) { V[nv][e] = w[e] * (1. / na); AV[nv][e] = aw[e] * (1. / na); }
            nv++;
            memcpy(q, qn, sizeof(double) * n);
            if (na < 1e-13) { restart = 1; break; }
        }
        for (long i = 0; i < nv; i++) { free(V[i]); free(AV[i]); }
        if (!restart || it >= maxit) break;
        restarts++;
    }
    A->apply(A, x, t);
    for (long e = 0; e < n; e++) t[e] = b[e] - t[e];
    out->iterations = (int)it;
    out->restarts = restarts;
    out->resnorm = nrm(n, t);
    free(V); free(AV); free(r); free(Mr); free(t); free(q); free(qn); free(w); free(aw);
    return 0;
}

static void make_ops(orc_op *A, csr_ctx *ac, long n, const long *rowptr, const int *cols, const double *vals)
{
    ac->n = n; ac->rowptr = rowptr; ac->cols = cols; ac->vals = vals;
    A->n = n; A->apply = csr_apply; A->ctx = ac;
}

int orc_solve(int method, long n, const long *rowptr, const int *cols, const double *vals,
              int precond, const orc_mg *mg, const long *mrow, const int *mcol, const double *mval,
              const double *b, double *x, double tol, long maxit, int restart, orc_solve_out *out, double *hist)
{
    orc_op A, M;
    csr_ctx ac, mc;
    make_ops(&A, &ac, n, rowptr, cols, vals);
    orc_op *Mp = NULL;
    if (precond == 1) { M.n = n; M.apply = ident_apply; M.ctx = NULL; Mp = &M; }
    else if (precond == 2) { M.n = n; M.apply = mg_apply_op; M.ctx = (void *)mg; Mp = &M; }
    else if (precond == 3) { make_ops(&M, &mc, n, mrow, mcol, mval); Mp = &M; }
    orc_result r;
    memset(&r, 0, sizeof(r));
    int rc = 0;
    switch (method) {
    case 0: rc = Mp ? orc_pcg(&A, Mp, b, x, tol, maxit, &r, hist) : orc_cg(&A, b, x, tol, maxit, &r, hist); break;
    case 1: rc = orc_bicgstab(&A, Mp, b, x, tol, &r); break;
    case 2: rc = orc_gmres(&A, Mp, b, x, tol, restart, (int)maxit, &r); break;
    case 3: rc = orc_gmres2(&A, precond == 1 ? NULL : Mp, b, x, tol, maxit, &r); break;
    default: return -1;
    }
    out->iterations = r.iterations; out->fell_back = r.fell_back; out->restarts = r.restarts; out->resnorm = r.resnorm;
    return rc;
}

/* ------------------------------------------------------------------------- */
/* SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number       */
/* generators", OOPSLA 2014) and Box-Muller normals for seeded synthetic input. */
/* ------------------------------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
