/*
 * cpu_baseline.c -- the CPU baseline timed beside the GPU: an OpenMP C restatement of the
 * reference's CPU solve path for tensor-product Qk spaces.  Test and benchmark infrastructure
 * like the rest of oracle/: bench.py's cpu_baseline leg and tests/ load it, the product never does.
 *
 * What it restates (reference paths relative to Core/src/main/java/):
 *  - the row-parallel CSR product of SparseMvMul.mvMul (linalg/SparseMvMul.java:75-94) for the
 *    level systems and the prolongation, and the column-parallel SparseMvMul.tvMul
 *    (:96-116) for the restriction, here as the row-parallel product with the transpose of P;
 *  - mgStepV (multigrid/MGPreconditionerInterface.java:125-198): smooth, restrict the defect
 *    with zero boundary values, recurse from zero, prolongate with zero boundary values and
 *    add, smooth; the coarsest level solved directly (:133-137; factorised once here);
 *  - the prolongation matrix of MGPreconditionerSpace (multigrid/MGPreconditionerSpace.java:
 *    59-76): P[g, f] = phi_f(node_g), built from the Lagrange basis of the coarse cell that
 *    contains node g, exact zeros skipped;
 *  - solvePCG (linalg/IterativeSolver.java:102-162) with x0 = 0 and an absolute tolerance;
 *  - the Chebyshev-Jacobi smoother of this project (not in the reference; Saad, Iterative
 *    Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1) on [lmax / ratio, lmax] of D^-1 A,
 *    lmax = safety * (power iteration estimate), the same recurrence as oracle/jfem_oracle.c.
 * Level matrices are passed in (borrowed host CSR arrays); every loop over rows or vector
 * entries runs on all OpenMP threads, as the reference's parallel streams run on the common
 * ForkJoin pool.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define CB_MAXD 3
#define CB_MAXLEV 16

typedef struct {
    long n;
    const long *rp;
    const int *ci;
    const double *v;
} cb_csr;

typedef struct {
    long n, ncols;
    long *rp;
    int *ci;
    double *v;
} cb_own;

typedef struct {
    int nlev, d, k, degree, runs;
    int cells[CB_MAXLEV][CB_MAXD];
    cb_csr A[CB_MAXLEV];
    double *dinv[CB_MAXLEV];
    unsigned char *bnd[CB_MAXLEV];
    cb_own P[CB_MAXLEV], PT[CB_MAXLEV];   /* between level l and l + 1: P fine x coarse */
    double lmin[CB_MAXLEV], lmax[CB_MAXLEV];
    double *x[CB_MAXLEV], *b[CB_MAXLEV], *r[CB_MAXLEV], *t[CB_MAXLEV], *dd[CB_MAXLEV];
    double *lu;   /* coarsest system, LU factors (row-major) */
    int *piv;
} cb_mg;

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int cb_threads(void) { return omp_get_max_threads(); }

/* ---------------------------------------------------------------- BLAS-1, SpMV */
static void spmv(const cb_csr *A, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < A->n; i++) {
        double s = 0;
        for (long p = A->rp[i]; p < A->rp[i + 1]; p++) s += A->v[p] * x[A->ci[p]];
        y[i] = s;
    }
}
static void spmv_own(const cb_own *A, const double *x, double *y)
{
    cb_csr B = {A->n, A->rp, A->ci, A->v};
    spmv(&B, x, y);
}
static double dot(long n, const double *x, const double *y)
{
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

/* ------------------------------------------ structured Qk space: DOF numbering */
/* the reference's first-appearance numbering (tensorproduct/ContinuousTPFESpace.java:29-50)
 * in closed form: cells in IntCoordinates order (last axis fastest), local nodes with axis 0
 * fastest, a node new in cell c iff l_a >= 1 on every axis with c_a > 0 */
static long dof_of_lattice(int d, int k, const int *cells, const int *p)
{
    long tail[CB_MAXD], t = 1, pre = 0, mprod = 1, rank = 0;
    for (int a = d - 1; a >= 0; a--) { tail[a] = t; t *= (long)k * cells[a] + 1; }
    for (int a = 0; a < d; a++) {
        const int c = p[a] == 0 ? 0 : (p[a] - 1) / k;
        const int l = p[a] - c * k;
        const long m = c > 0 ? k : k + 1, S = c > 0 ? (long)c * k + 1 : 0;
        pre += mprod * S * tail[a];
        rank += (long)(l - (c > 0 ? 1 : 0)) * mprod;
        mprod *= m;
    }
    return pre + rank;
}

static long ndofs(int d, int k, const int *cells)
{
    long n = 1;
    for (int a = 0; a < d; a++) n *= (long)k * cells[a] + 1;
    return n;
}

/* 1D prolongation row of fine lattice node i (fine grid 2 nc cells): coarse lattice columns
 * and Lagrange weights of the coarse cell containing the node, exact zeros skipped */
static int prol_row_1d(int k, int nc, int i, int *col, double *w)
{
    int c = i / (2 * k);
    if (c >= nc) c = nc - 1;
    const double tt = (double)(i - c * 2 * k) / (2.0 * k);   /* position in the coarse cell */
    int m = 0;
    for (int j = 0; j <= k; j++) {
        double v = 1.0;
        for (int q = 0; q <= k; q++)
            if (q != j) v *= (tt - (double)q / k) / ((double)(j - q) / k);
        if (v != 0.0) { col[m] = c * k + j; w[m] = v; m++; }
    }
    return m;
}

static void build_prolongation(cb_mg *G, int l)
{
    const int d = G->d, k = G->k;
    const int *cc = G->cells[l], *cf = G->cells[l + 1];
    long nf = ndofs(d, k, cf), nc = ndofs(d, k, cc);
    cb_own *P = &G->P[l];
    P->n = nf; P->ncols = nc;
    /* per-axis 1D rows */
    int *c1[CB_MAXD], *n1[CB_MAXD];
    double *w1[CB_MAXD];
    long len[CB_MAXD];
    for (int a = 0; a < d; a++) {
        len[a] = (long)k * cf[a] + 1;
        c1[a] = malloc(sizeof(int) * len[a] * (k + 1));
        w1[a] = malloc(sizeof(double) * len[a] * (k + 1));
        n1[a] = malloc(sizeof(int) * len[a]);
        for (long i = 0; i < len[a]; i++)
            n1[a][i] = prol_row_1d(k, cc[a], (int)i, c1[a] + i * (k + 1), w1[a] + i * (k + 1));
    }
    /* fine DOF of every fine lattice point (last axis fastest) */
    long *lat2dof = malloc(sizeof(long) * nf);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nf; q++) {
        int p[CB_MAXD];
        long rem = q;
        for (int a = d - 1; a >= 0; a--) { p[a] = (int)(rem % len[a]); rem /= len[a]; }
        lat2dof[q] = dof_of_lattice(d, k, cf, p);
    }
    P->rp = calloc(nf + 1, sizeof(long));
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nf; q++) {
        long rem = q, cnt = 1;
        for (int a = d - 1; a >= 0; a--) { cnt *= n1[a][rem % len[a]]; rem /= len[a]; }
        P->rp[lat2dof[q] + 1] = cnt;
    }
    for (long i = 0; i < nf; i++) P->rp[i + 1] += P->rp[i];
    P->ci = malloc(sizeof(int) * P->rp[nf]);
    P->v = malloc(sizeof(double) * P->rp[nf]);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nf; q++) {
        int p[CB_MAXD], idx[CB_MAXD] = {0, 0, 0};
        long rem = q;
        for (int a = d - 1; a >= 0; a--) { p[a] = (int)(rem % len[a]); rem /= len[a]; }
        long o = P->rp[lat2dof[q]];
        for (;;) {
            int pc[CB_MAXD];
            double w = 1.0;
            for (int a = 0; a < d; a++) {
                pc[a] = c1[a][(long)p[a] * (k + 1) + idx[a]];
                w *= w1[a][(long)p[a] * (k + 1) + idx[a]];
            }
            P->ci[o] = (int)dof_of_lattice(d, k, cc, pc);
            P->v[o] = w;
            o++;
            int a = 0;
            while (a < d && ++idx[a] == n1[a][p[a]]) { idx[a] = 0; a++; }
            if (a == d) break;
        }
    }
    free(lat2dof);
    for (int a = 0; a < d; a++) { free(c1[a]); free(w1[a]); free(n1[a]); }
    /* transpose (restriction rows) */
    cb_own *T = &G->PT[l];
    T->n = nc; T->ncols = nf;
    T->rp = calloc(nc + 2, sizeof(long));
    for (long p = 0; p < P->rp[nf]; p++) T->rp[P->ci[p] + 2]++;
    for (long i = 0; i < nc; i++) T->rp[i + 2] += T->rp[i + 1];
    T->ci = malloc(sizeof(int) * P->rp[nf]);
    T->v = malloc(sizeof(double) * P->rp[nf]);
    for (long i = 0; i < nf; i++)
        for (long p = P->rp[i]; p < P->rp[i + 1]; p++) {
            const long o = T->rp[P->ci[p] + 1]++;
            T->ci[o] = (int)i;
            T->v[o] = P->v[p];
        }
}

/* ------------------------------------------------------------- smoother */
static void smooth(cb_mg *G, int l, const double *b, double *x)
{
    const long n = G->A[l].n;
    double *t = G->t[l], *dd = G->dd[l];
    const double *dinv = G->dinv[l];
    const double theta = 0.5 * (G->lmax[l] + G->lmin[l]), delta = 0.5 * (G->lmax[l] - G->lmin[l]);
    const double sigma = theta / delta;
    for (int app = 0; app < G->runs; app++) {
        double rho_old = 1.0 / sigma;
        for (int kd = 0; kd < G->degree; kd++) {
            spmv(&G->A[l], x, t);
            double c1 = 0, c2 = 0;
            if (kd > 0) {
                const double rho = 1.0 / (2.0 * sigma - rho_old);
                c1 = rho * rho_old;
                c2 = 2.0 * rho / delta;
                rho_old = rho;
            }
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; i++) {
                const double r = b[i] - t[i];
                const double v = kd == 0 ? (dinv[i] * r) / theta : c1 * dd[i] + c2 * (dinv[i] * r);
                dd[i] = v;
                x[i] = x[i] + v;
            }
        }
    }
}

static void coarse_solve(const cb_mg *G, const double *b, double *x)
{
    const long n = G->A[0].n;
    double *y = malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) y[i] = b[G->piv[i]];
    for (long i = 0; i < n; i++) {
        double s = y[i];
        for (long j = 0; j < i; j++) s -= G->lu[i * n + j] * y[j];
        y[i] = s;
    }
    for (long i = n - 1; i >= 0; i--) {
        double s = y[i];
        for (long j = i + 1; j < n; j++) s -= G->lu[i * n + j] * x[j];
        x[i] = s / G->lu[i * n + i];
    }
    free(y);
}

/* mgStepV(level, x, b), x updated in place */
static void step_v(cb_mg *G, int l, double *x, const double *b)
{
    if (l == 0) { coarse_solve(G, b, x); return; }
    const long n = G->A[l].n, nc = G->A[l - 1].n;
    double *r = G->r[l], *xc = G->x[l - 1], *bc = G->b[l - 1];
    const unsigned char *bf = G->bnd[l], *bcm = G->bnd[l - 1];
    smooth(G, l, b, x);
    spmv(&G->A[l], x, r);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) r[i] = b[i] - r[i];
    spmv_own(&G->PT[l - 1], r, bc);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nc; i++) {
        if (bcm[i]) bc[i] = 0;
        xc[i] = 0;
    }
    step_v(G, l - 1, xc, bc);
    spmv_own(&G->P[l - 1], xc, r);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) x[i] = x[i] + (bf[i] ? 0.0 : r[i]);
    smooth(G, l, b, x);
}

/* ---------------------------------------------------------------- set-up */
static double power_lmax(const cb_csr *A, const double *dinv, int its)
{
    const long n = A->n;
    double *v = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        v[i] = 0.5 + (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
    double nv = sqrt(dot(n, v, v)), lam = 0;
    for (long i = 0; i < n; i++) v[i] /= nv;
    for (int it = 0; it < its; it++) {
        spmv(A, v, w);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) w[i] *= dinv[i];
        lam = sqrt(dot(n, w, w));
        if (lam == 0) break;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) v[i] = w[i] / lam;
    }
    free(v); free(w);
    return lam;
}

static void boundary_mask(int d, int k, const int *cells, unsigned char *mask)
{
    long len[CB_MAXD];
    long n = 1;
    for (int a = 0; a < d; a++) { len[a] = (long)k * cells[a] + 1; n *= len[a]; }
#pragma omp parallel for schedule(static)
    for (long q = 0; q < n; q++) {
        int p[CB_MAXD], on = 0;
        long rem = q;
        for (int a = d - 1; a >= 0; a--) {
            p[a] = (int)(rem % len[a]);
            rem /= len[a];
            if (p[a] == 0 || p[a] == len[a] - 1) on = 1;
        }
        mask[dof_of_lattice(d, k, cells, p)] = (unsigned char)on;
    }
}

/* levels 0 (coarsest, coarse_cells) .. nlev-1, each with twice the cells of the one before;
 * n/rp/ci/v: the level systems after Dirichlet elimination (borrowed, kept alive by the caller) */
void *cb_mg_create(int d, int k, const int *coarse_cells, int nlev, const long *n, const long *const *rp,
                   const int *const *ci, const double *const *v, int degree, int runs, double ratio, double safety,
                   int power_its)
{
    if (d < 1 || d > CB_MAXD || nlev < 1 || nlev > CB_MAXLEV) return NULL;
    cb_mg *G = calloc(1, sizeof(cb_mg));
    G->nlev = nlev; G->d = d; G->k = k; G->degree = degree; G->runs = runs;
    for (int l = 0; l < nlev; l++) {
        for (int a = 0; a < d; a++) G->cells[l][a] = coarse_cells[a] << l;
        if (ndofs(d, k, G->cells[l]) != n[l]) { free(G); return NULL; }
        G->A[l].n = n[l]; G->A[l].rp = rp[l]; G->A[l].ci = ci[l]; G->A[l].v = v[l];
        G->dinv[l] = malloc(sizeof(double) * n[l]);
        const cb_csr *A = &G->A[l];
        double *dinv = G->dinv[l];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n[l]; i++) {           /* JacobiSmoother ctor (JacobiSmoother.java:14-25) */
            double dg = 0;
            for (long p = A->rp[i]; p < A->rp[i + 1]; p++) if (A->ci[p] == i) dg += A->v[p];
            if (fabs(dg) < 1e-10) dg = 1;
            dinv[i] = 1. / dg;
        }
        G->bnd[l] = malloc(n[l]);
        boundary_mask(d, k, G->cells[l], G->bnd[l]);
        G->x[l] = calloc(n[l], sizeof(double));
        G->b[l] = calloc(n[l], sizeof(double));
        G->r[l] = calloc(n[l], sizeof(double));
        G->t[l] = calloc(n[l], sizeof(double));
        G->dd[l] = calloc(n[l], sizeof(double));
        if (l > 0) {
            G->lmax[l] = safety * power_lmax(A, dinv, power_its);
            G->lmin[l] = G->lmax[l] / ratio;
        }
    }
    for (int l = 0; l + 1 < nlev; l++) build_prolongation(G, l);
    /* coarsest level: dense LU with partial pivoting, factorised once */
    const long n0 = n[0];
    G->lu = calloc(n0 * n0, sizeof(double));
    G->piv = malloc(sizeof(int) * n0);
    for (long i = 0; i < n0; i++)
        for (long p = rp[0][i]; p < rp[0][i + 1]; p++) G->lu[i * n0 + ci[0][p]] += v[0][p];
    for (long i = 0; i < n0; i++) G->piv[i] = (int)i;
    for (long c = 0; c < n0; c++) {
        long pr = c;
        for (long i = c + 1; i < n0; i++) if (fabs(G->lu[i * n0 + c]) > fabs(G->lu[pr * n0 + c])) pr = i;
        if (pr != c) {
            for (long j = 0; j < n0; j++) {
                double tmp = G->lu[c * n0 + j];
                G->lu[c * n0 + j] = G->lu[pr * n0 + j];
                G->lu[pr * n0 + j] = tmp;
            }
            int tp = G->piv[c]; G->piv[c] = G->piv[pr]; G->piv[pr] = tp;
        }
        for (long i = c + 1; i < n0; i++) {
            const double f = G->lu[i * n0 + c] / G->lu[c * n0 + c];
            G->lu[i * n0 + c] = f;
            for (long j = c + 1; j < n0; j++) G->lu[i * n0 + j] -= f * G->lu[c * n0 + j];
        }
    }
    return G;
}

void cb_mg_free(void *p)
{
    cb_mg *G = p;
    if (!G) return;
    for (int l = 0; l < G->nlev; l++) {
        free(G->dinv[l]); free(G->bnd[l]);
        free(G->x[l]); free(G->b[l]); free(G->r[l]); free(G->t[l]); free(G->dd[l]);
        if (l + 1 < G->nlev) {
            free(G->P[l].rp); free(G->P[l].ci); free(G->P[l].v);
            free(G->PT[l].rp); free(G->PT[l].ci); free(G->PT[l].v);
        }
    }
    free(G->lu); free(G->piv); free(G);
}

/* the prolongation between level l and l + 1 as CSR (fine rows), for the tests */
long cb_mg_prol_nnz(const void *p, int l) { const cb_mg *G = p; return G->P[l].rp[G->P[l].n]; }
void cb_mg_prol_csr(const void *p, int l, long *rp, int *ci, double *v)
{
    const cb_mg *G = p;
    const cb_own *P = &G->P[l];
    memcpy(rp, P->rp, sizeof(long) * (P->n + 1));
    memcpy(ci, P->ci, sizeof(int) * P->rp[P->n]);
    memcpy(v, P->v, sizeof(double) * P->rp[P->n]);
}
double cb_mg_lmax(const void *p, int l) { return ((const cb_mg *)p)->lmax[l]; }

/* one V-cycle from x (in/out) on the finest level */
void cb_mg_vcycle(void *p, double *x, const double *b)
{
    cb_mg *G = p;
    step_v(G, G->nlev - 1, x, b);
}

/* solvePCG with the V-cycle preconditioner (one V-cycle from zero per application), x0 = x on
 * entry, stop when ||r||_2 <= tol (absolute) or after maxit iterations; returns the iteration
 * count, *seconds the wall time of the solve, *resnorm the final recursive residual norm */
int cb_mg_pcg(void *p, const double *b, double *x, double tol, long maxit, double *seconds, double *resnorm)
{
    cb_mg *G = p;
    const int L = G->nlev - 1;
    const long n = G->A[L].n;
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *pp = malloc(sizeof(double) * n),
           *q = malloc(sizeof(double) * n);
    const double t0 = now();
    spmv(&G->A[L], x, q);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) { r[i] = b[i] - q[i]; z[i] = 0; }
    step_v(G, L, z, r);
    memcpy(pp, z, sizeof(double) * n);
    double rz = dot(n, r, z), rr = dot(n, r, r);
    int it = 0;
    for (; it < maxit && sqrt(rr) > tol; it++) {
        spmv(&G->A[L], pp, q);
        const double alpha = rz / dot(n, pp, q);
        double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (long i = 0; i < n; i++) {
            x[i] += alpha * pp[i];
            r[i] -= alpha * q[i];
            s += r[i] * r[i];
            z[i] = 0;
        }
        rr = s;
        step_v(G, L, z, r);
        const double rz2 = dot(n, r, z), beta = rz2 / rz;
        rz = rz2;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) pp[i] = z[i] + beta * pp[i];
    }
    *seconds = now() - t0;
    *resnorm = sqrt(rr);
    free(r); free(z); free(pp); free(q);
    return it;
}

/* mean seconds of the finest-level row-parallel SpMV over reps products */
double cb_mg_spmv_seconds(void *p, int reps)
{
    cb_mg *G = p;
    const int L = G->nlev - 1;
    const long n = G->A[L].n;
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) x[i] = 1.0 / (1.0 + (double)(i % 97));
    spmv(&G->A[L], x, y);
    const double t0 = now();
    for (int r = 0; r < reps; r++) spmv(&G->A[L], x, y);
    const double s = (now() - t0) / reps;
    free(x); free(y);
    return s;
}
