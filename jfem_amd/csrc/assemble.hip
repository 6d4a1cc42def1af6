// Structured Qk space (ContinuousTPFESpace) and CSR assembly of TPCellIntegral
// GRAD_GRAD / VALUE_VALUE / H1 on gfx950.
//
// Element matrices: the reference evaluates every element entry by tensor
// Gauss quadrature in a fixed floating-point order
// (TPCellIntegral.integrateNonTensorProduct, M/tensorproduct/TPCellIntegral.java:78-102,
// with ContinuousTPShapeFunction.fastGradientInCell :173-198).  That value only
// depends, per axis, on the 1D cell's quadrature weights and basis tables, so
// 1D cells with bit-identical tables form a class and one element matrix per
// class tuple reproduces the reference's arithmetic for every cell of the tuple
// (computed on the GPU with contraction disabled, operation for operation).
//
// CSR: one group of lanes per row (row-owner gather, deterministic, no
// atomics).  The columns of row i are the lattice points of the cells around
// node i; in the reference's first-appearance numbering their ascending order
// is: owner cell in lexicographic order, then rank inside the owner cell, so
// the sorted column list is enumerated directly without sorting.  An entry
// exists iff some cell contribution is nonzero (MatrixFESpace.addToMatrix
// skips exact zeros, M/basic/MatrixFESpace.java:115-122) or it is the dummy
// (0,0) entry of a new SparseMatrix (M/linalg/SparseMatrix.java:95-102);
// values are summed over cells in lexicographic cell order.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

namespace jfem {

// ------------------------------- host tables -------------------------------

// QuadratureRule1D (M/tensorproduct/QuadratureRule1D.java:11-92)
#pragma clang fp contract(off)
static int quad_rule(int rule, double* pts, double* wts)
{
    double p[6], w[6];
    int n = 0;
    switch (rule) {
    case 1: n = 1; w[0] = 2; p[0] = 0; break;
    case 2: n = 2; w[0] = 1; w[1] = 1; p[0] = -std::sqrt(1. / 3); p[1] = std::sqrt(1. / 3); break;
    case 3: n = 3; w[0] = 5. / 9; w[1] = 8. / 9; w[2] = 5. / 9;
        p[0] = -std::sqrt(3. / 5); p[1] = 0; p[2] = std::sqrt(3. / 5); break;
    case 4: n = 4;
        w[0] = (18. - std::sqrt(30.)) / 36; w[1] = (18. + std::sqrt(30.)) / 36;
        w[2] = (18. + std::sqrt(30.)) / 36; w[3] = (18. - std::sqrt(30.)) / 36;
        p[0] = -std::sqrt(3. / 7 + 2. / 7 * std::sqrt(6. / 5)); p[1] = -std::sqrt(3. / 7 - 2. / 7 * std::sqrt(6. / 5));
        p[2] = std::sqrt(3. / 7 - 2. / 7 * std::sqrt(6. / 5));  p[3] = std::sqrt(3. / 7 + 2. / 7 * std::sqrt(6. / 5));
        break;
    case 5: n = 5;
        w[0] = (322.0 - 13 * std::sqrt(70.)) / 900; w[1] = (322.0 + 13 * std::sqrt(70.)) / 900;
        w[2] = 128.0 / 225;
        w[3] = (322.0 + 13 * std::sqrt(70.)) / 900; w[4] = (322.0 - 13 * std::sqrt(70.)) / 900;
        p[0] = -std::sqrt(5.0 + 2 * std::sqrt(10.0 / 7)) / 3; p[1] = -std::sqrt(5.0 - 2 * std::sqrt(10.0 / 7)) / 3;
        p[2] = 0;
        p[3] = std::sqrt(5.0 - 2 * std::sqrt(10.0 / 7)) / 3;  p[4] = std::sqrt(5.0 + 2 * std::sqrt(10.0 / 7)) / 3;
        break;
    case 12: n = 2; w[0] = 1. / 2; w[1] = 1. / 2; p[0] = -1; p[1] = 1; break;
    case 13: n = 3; w[0] = 1. / 3; w[1] = 4. / 3; w[2] = 1. / 3; p[0] = -1; p[1] = 0; p[2] = 1; break;
    case 14: n = 4; w[0] = 1. / 6; w[1] = 5. / 6; w[2] = 5. / 6; w[3] = 1. / 6;
        p[0] = -1; p[1] = -std::sqrt(1. / 5); p[2] = std::sqrt(1. / 5); p[3] = 1; break;
    case 15: n = 5; w[0] = 1. / 10; w[1] = 49. / 90; w[2] = 32. / 45; w[3] = 49. / 90; w[4] = 1. / 10;
        p[0] = -1; p[1] = -std::sqrt(3. / 7); p[2] = 0; p[3] = std::sqrt(3. / 7); p[4] = 1; break;
    case 16: n = 6;
        w[0] = 1. / 15; w[1] = (14. - std::sqrt(7.)) / 30; w[2] = (14. + std::sqrt(7.)) / 30;
        w[3] = (14. + std::sqrt(7.)) / 30; w[4] = (14. - std::sqrt(7.)) / 30; w[5] = 1. / 15;
        p[0] = -1; p[1] = -std::sqrt(1. / 3 + 2 * std::sqrt(7.) / 21); p[2] = -std::sqrt(1. / 3 - 2 * std::sqrt(7.) / 21);
        p[3] = +std::sqrt(1. / 3 - 2 * std::sqrt(7.) / 21); p[4] = +std::sqrt(1. / 3 + 2 * std::sqrt(7.) / 21); p[5] = 1;
        break;
    default: return -1;
    }
    for (int i = 0; i < n; i++) { pts[i] = p[i] / 2 + 0.5; wts[i] = w[i] / 2; }
    return n;
}

// LagrangeBasisFunction1D.valueOnReferenceCell / derivativeOnReferenceCell
// (M/tensorproduct/LagrangeBasisFunction1D.java:55-154)
__host__ __device__ static double lag_value(double pos, int l, int k)
{
    if (k == 1) return l == 0 ? 1.0 - pos : pos;
    if (k == 2) {
        if (l == 0) return 1.0 - 3.0 * pos + 2.0 * pos * pos;
        if (l == 1) return (pos - pos * pos) * 4.0;
        return 2.0 * pos * pos - pos;
    }
    double ret = 1;
    for (int i = 0; i <= k; i++)
        if (i != l) ret *= (pos - 1. * i / k) / (1. * l / k - 1. * i / k);
    return ret;
}

__host__ __device__ static double lag_deriv(double pos, int l, int k)
{
    if (k == 1) return l == 0 ? -1.0 : 1.0;
    if (k == 2) {
        if (l == 0) return -3.0 + 4.0 * pos;
        if (l == 1) return (1.0 - 2.0 * pos) * 4.0;
        return 4.0 * pos - 1.0;
    }
    double derret = 0;
    for (int j = 0; j <= k; j++) {
        if (j == l) continue;
        double ret = 1. / (1. * l / k - 1. * j / k);
        for (int i = 0; i <= k; i++)
            if (i != l && i != j) ret *= (pos - 1. * i / k) / (1. * l / k - 1. * i / k);
        derret += ret;
    }
    return derret;
}

// Cell1D.isInCell with PerformanceArguments.doubleTolerance = 1e-10 (Cell1D.java:55-59)
__host__ __device__ static bool in_cell(double s, double e, double pos) { return pos >= s - 100 * 1e-10 && pos <= e + 100 * 1e-10; }
#pragma clang fp contract(on)

// 1D table of one cell: weights w[q] = refw*len, value[l][q], deriv[l][q]
constexpr int MAXQ = 6;
constexpr int MAXK1 = 5;   // k <= 4
struct Tab1D {
    double w[MAXQ];
    double val[MAXK1][MAXQ];
    double der[MAXK1][MAXQ];
};

#pragma clang fp contract(off)
static Tab1D make_tab(double s, double e, int k, const double* qp, const double* qw, int Q)
{
    Tab1D t;
    std::memset(&t, 0, sizeof(t));
    const double len = e - s;
    for (int q = 0; q < Q; q++) {
        const double pos = qp[q] * len + s;   // Cell1D.distributeQuadrature :29-30
        t.w[q] = qw[q] * len;
        for (int l = 0; l <= k; l++) {
            t.val[l][q] = in_cell(s, e, pos) ? lag_value((pos - s) / len, l, k) : 0;
            t.der[l][q] = in_cell(s, e, pos) ? lag_deriv((pos - s) / len, l, k) / len : 0;
        }
    }
    return t;
}
#pragma clang fp contract(on)

// ------------------------------- kernels ----------------------------------

struct ElemParams {
    int d, k, Q, nloc, nint;
    int kinds[4];
    double wconst[4];
    double wvec[4][MAXD];   // constant vector weight of GRAD_VALUE / VALUE_GRAD
    int full;               // 1: every (iv, iu) pair is listed (non-symmetric integrals), no mirrored write
    int ncls[MAXD];
    long ntuples;
};

// E[(i*ntuples + t)*nloc*nloc + iv*nloc + iu] = integral i on class tuple t
// between u = shapeFunction1 (local iu) and v = shapeFunction2 (local iv).
// Operation order follows integrateNonTensorProduct exactly (no contraction).
// One thread per (integral, class tuple, node pair iv <= iu): the value is
// written to E[iv][iu] and E[iu][iv].  Swapping the two functions leaves every
// operation of the sum unchanged (products commute exactly in IEEE arithmetic
// and the quadrature order is the same), so the table is symmetric bit for
// bit.  With a non-symmetric integral (GRAD_VALUE, VALUE_GRAD) every pair is
// listed (P.full) and only E[iv][iu] is written.  DD, KK, QQ > 0 fix dimension, degree and points per axis, so the
// quadrature loop unrolls and the 1D tables stay in registers.
template <int DD, int KK, int QQ>
__global__ __launch_bounds__(256) void k_element_tables(ElemParams P, const Tab1D* __restrict__ tabs0,
                                                        const Tab1D* __restrict__ tabs1,
                                                        const Tab1D* __restrict__ tabs2,
                                                        const int* __restrict__ pairs, double* __restrict__ E)
{
#pragma clang fp contract(off)
    const int d = DD ? DD : P.d, k = KK ? KK : P.k, Q = QQ ? QQ : P.Q;
    const long nent = (long)P.nloc * P.nloc;
    const long npair = P.full ? (long)P.nloc * P.nloc : (long)P.nloc * (P.nloc + 1) / 2;
    const long total = npair * P.ntuples * P.nint;
    for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
        const int pr = pairs[id % npair];
        const int iv = pr >> 16, iu = pr & 0xffff;
        const long t = (id / npair) % P.ntuples;
        const int ii = (int)(id / (npair * P.ntuples));
        int cls[MAXD] = {0, 0, 0};
        long r = t;
        for (int a = d - 1; a >= 0; a--) { cls[a] = (int)(r % P.ncls[a]); r /= P.ncls[a]; }
        const Tab1D* tb[MAXD] = {tabs0 + cls[0], d > 1 ? tabs1 + cls[1] : tabs0, d > 2 ? tabs2 + cls[2] : tabs0};
        int lu[MAXD] = {0, 0, 0}, lv[MAXD] = {0, 0, 0};
        int ru = iu, rv = iv;
#pragma unroll
        for (int a = 0; a < MAXD; a++)
            if (a < d) { lu[a] = ru % (k + 1); ru /= k + 1; lv[a] = rv % (k + 1); rv /= k + 1; }
        const int kind = P.kinds[ii];
        const double wc = P.wconst[ii];
        int nq = 1;
        for (int a = 0; a < d; a++) nq *= Q;
        // quadrature points of the cell, a constant with fixed d and Q (1: the loop is not unrolled)
        constexpr int QN = QQ && DD ? (DD == 3 ? QQ * QQ * QQ : DD == 2 ? QQ * QQ : QQ) : 1;
        double ret = 0;
        if constexpr (DD > 0 && KK > 0 && QQ > 0) {
            // fixed shape: the quadrature sum below is the generic one, operation for operation, with
            // the 1D values, derivatives and weights of axes 1.. held in registers and those of axis
            // 0 loaded once per axis-0 point (the points are visited in the same order, last axis
            // fastest)
            constexpr int R = DD > 1 ? QN / QQ : 1;   // points of axes 1.. per axis-0 point
            double Vu[MAXD][QQ], Vv[MAXD][QQ], Du[MAXD][QQ], Dv[MAXD][QQ], Wq[MAXD][QQ];
#pragma unroll
            for (int a = 1; a < DD; a++)
#pragma unroll
                for (int c = 0; c < QQ; c++) {
                    Vu[a][c] = tb[a]->val[lu[a]][c];
                    Vv[a][c] = tb[a]->val[lv[a]][c];
                    Du[a][c] = tb[a]->der[lu[a]][c];
                    Dv[a][c] = tb[a]->der[lv[a]][c];
                    Wq[a][c] = tb[a]->w[c];
                }
            const double* wv = P.wvec[ii];
#pragma unroll 1
            for (int c0 = 0; c0 < QQ; c0++) {
                const double vu0 = tb[0]->val[lu[0]][c0], vv0 = tb[0]->val[lv[0]][c0];
                const double du0 = tb[0]->der[lu[0]][c0], dv0 = tb[0]->der[lv[0]][c0], w0 = tb[0]->w[c0];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    // this point's 1D factors per axis
                    double pu[DD], pv[DD], qu[DD], qv[DD], pw[DD];
                    pu[0] = vu0; pv[0] = vv0; qu[0] = du0; qv[0] = dv0; pw[0] = w0;
                    int rr = r;
#pragma unroll
                    for (int a = DD - 1; a >= 1; a--) {
                        const int c = rr % QQ;
                        rr /= QQ;
                        pu[a] = Vu[a][c]; pv[a] = Vv[a][c]; qu[a] = Du[a][c]; qv[a] = Dv[a][c]; pw[a] = Wq[a][c];
                    }
                    double val;
                    double vu = 1, vv = 1;
                    if (kind != 0) {
#pragma unroll
                        for (int a = 0; a < DD; a++) { vu *= pu[a]; vv *= pv[a]; }
                    }
                    if (kind == 1) {
                        val = vu * vv * wc;
                    } else if (kind >= 3) {
                        double inner = 0;
#pragma unroll
                        for (int i = 0; i < DD; i++) {
                            double gg = 1;
#pragma unroll
                            for (int j = 0; j < DD; j++)
                                gg *= i == j ? (kind == 3 ? qu[j] : qv[j]) : (kind == 3 ? pu[j] : pv[j]);
                            inner += gg * wv[i];
                        }
                        val = inner * (kind == 3 ? vv : vu);
                    } else {
                        double inner = 0;
#pragma unroll
                        for (int i = 0; i < DD; i++) {
                            double g1 = 1, g2 = 1;
#pragma unroll
                            for (int j = 0; j < DD; j++) {
                                if (i == j) { g1 *= qu[j]; g2 *= qv[j]; }
                                else { g1 *= pu[j]; g2 *= pv[j]; }
                            }
                            inner += g1 * g2;
                        }
                        val = kind == 0 ? inner * wc : (vu * vv + inner) * wc;
                    }
#pragma unroll
                    for (int a = 0; a < DD; a++) val *= pw[a];
                    ret += val;
                }
            }
            const long base = ((long)ii * P.ntuples + t) * nent;
            E[base + (long)iv * P.nloc + iu] = ret;
            if (!P.full) E[base + (long)iu * P.nloc + iv] = ret;
            continue;
        }
#pragma unroll QN
        for (int q = 0; q < (QN > 1 ? QN : nq); q++) {
            int c[MAXD] = {0, 0, 0};
            int qq = q;
#pragma unroll
            for (int a = MAXD - 1; a >= 0; a--)
                if (a < d) { c[a] = qq % Q; qq /= Q; }   // last axis fastest
            double val;
            double vu = 1, vv = 1;
            if (kind != 0) {
#pragma unroll
                for (int a = 0; a < MAXD; a++)
                    if (a < d) { vu *= tb[a]->val[lu[a]][c[a]]; vv *= tb[a]->val[lv[a]][c[a]]; }
            }
            if (kind == 1) {
                val = vu * vv * wc;
            } else if (kind >= 3) {
                // GRAD_VALUE: (grad u . w) v; VALUE_GRAD: (grad v . w) u (TPCellIntegral.java:142-157)
                const int* lg = kind == 3 ? lu : lv;
                double inner = 0;
#pragma unroll
                for (int i = 0; i < MAXD; i++) {
                    if (i >= d) break;
                    double g = 1;
#pragma unroll
                    for (int j = 0; j < MAXD; j++)
                        if (j < d) g *= i == j ? tb[j]->der[lg[j]][c[j]] : tb[j]->val[lg[j]][c[j]];
                    inner += g * P.wvec[ii][i];
                }
                val = inner * (kind == 3 ? vv : vu);
            } else {
                double inner = 0;
#pragma unroll
                for (int i = 0; i < MAXD; i++) {
                    if (i >= d) break;
                    double g1 = 1, g2 = 1;
#pragma unroll
                    for (int j = 0; j < MAXD; j++)
                        if (j < d) {
                            if (i == j) { g1 *= tb[j]->der[lu[j]][c[j]]; g2 *= tb[j]->der[lv[j]][c[j]]; }
                            else { g1 *= tb[j]->val[lu[j]][c[j]]; g2 *= tb[j]->val[lv[j]][c[j]]; }
                        }
                    inner += g1 * g2;
                }
                val = kind == 0 ? inner * wc : (vu * vv + inner) * wc;
            }
#pragma unroll
            for (int a = 0; a < MAXD; a++)
                if (a < d) val *= tb[a]->w[c[a]];
            ret += val;
        }
        const long base = ((long)ii * P.ntuples + t) * nent;
        E[base + (long)iv * P.nloc + iu] = ret;
        if (!P.full) E[base + (long)iu * P.nloc + iv] = ret;
    }
}

struct RowParams {
    GridDev g;
    int nloc, nint;
    long ntuples;
    const int* cls[MAXD];    // class id of each 1D cell per axis
    int ncls[MAXD];
    const double* E;         // element tables
    const double* coeff;     // per-cell coefficient or null
};

// per-axis candidate segments of a row: owner cell, first lattice point, count
struct AxisSegs {
    int nseg;
    int owner[3], first[3], cnt[3];
    int W;              // total lattice points
    int clo, chi;       // cells containing the row's node
};

// The row kernel is instantiated for fixed (dimension, degree) pairs, so that
// index loops unroll and divisions by the degree become constant divisions;
// DD = KK = 0 is the generic instantiation that reads them from the grid.
template <int DD, int KK>
struct DK {
    __device__ static int d(const GridDev& g) { return DD ? DD : g.d; }
    __device__ static int k(const GridDev& g) { return KK ? KK : g.k; }
};

// DOF index -> lattice point (the inverse of dof_of_lat_t).  All quantities
// are below the DOF count (< 2^31), so 32-bit unsigned arithmetic is exact.
template <int DD, int KK>
__device__ inline void lat_of_dof_t(const GridDev& g, long dof, int* p)
{
    const int d = DK<DD, KK>::d(g), k = DK<DD, KK>::k(g);
    unsigned tail[MAXD] = {1, 1, 1};
    unsigned t = 1;
#pragma unroll
    for (int a = MAXD - 1; a >= 0; a--)
        if (a < d) { tail[a] = t; t *= (unsigned)(k * g.n[a] + 1); }
    unsigned rem = (unsigned)dof, mprod = 1;
    int c[MAXD] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < MAXD; a++)
        if (a < d) {
            const unsigned unit = mprod * tail[a];
            const unsigned q = rem / unit;
            c[a] = q < (unsigned)(k + 1) ? 0 : (int)((q - 1) / (unsigned)k);
            rem -= (c[a] > 0 ? (unsigned)(c[a] * k + 1) : 0u) * unit;
            mprod *= c[a] > 0 ? k : k + 1;
        }
#pragma unroll
    for (int a = 0; a < MAXD; a++)
        if (a < d) {
            const unsigned m = c[a] > 0 ? k : k + 1;
            p[a] = c[a] * k + (int)(rem % m) + (c[a] > 0 ? 1 : 0);
            rem /= m;
        }
}

// lattice points per axis product of the later axes: tail[a] = prod_{b > a} (k n_b + 1)
template <int DD, int KK>
__device__ inline void lattice_tails(const GridDev& g, long* tail)
{
    const int d = DK<DD, KK>::d(g), k = DK<DD, KK>::k(g);
    long t = 1;
#pragma unroll
    for (int a = MAXD - 1; a >= 0; a--)
        if (a < d) { tail[a] = t; t *= (long)k * g.n[a] + 1; }
}

template <int DD, int KK>
__device__ inline long dof_of_lat_tails(const GridDev& g, const long* tail, const int* p)
{
    // every partial sum is at most the DOF index (< 2^31): 32-bit arithmetic
    const int d = DK<DD, KK>::d(g), k = DK<DD, KK>::k(g);
    unsigned pre = 0, mprod = 1, rank = 0;
#pragma unroll
    for (int a = 0; a < MAXD; a++)
        if (a < d) {
            const int c = p[a] == 0 ? 0 : (p[a] - 1) / k;
            const int l = p[a] - c * k;
            pre += mprod * (c > 0 ? (unsigned)(c * k + 1) : 0u) * (unsigned)tail[a];
            rank += (unsigned)(l - (c > 0 ? 1 : 0)) * mprod;
            mprod *= c > 0 ? k : k + 1;
        }
    return (long)(pre + rank);
}

template <int DD, int KK>
__device__ inline long dof_of_lat_t(const GridDev& g, const int* p)
{
    const int d = DK<DD, KK>::d(g), k = DK<DD, KK>::k(g);
    long tail[MAXD];
    long t = 1;
#pragma unroll
    for (int a = MAXD - 1; a >= 0; a--)
        if (a < d) { tail[a] = t; t *= (long)k * g.n[a] + 1; }
    long pre = 0, mprod = 1, rank = 0;
#pragma unroll
    for (int a = 0; a < MAXD; a++)
        if (a < d) {
            const int c = p[a] == 0 ? 0 : (p[a] - 1) / k;
            const int l = p[a] - c * k;
            pre += mprod * (c > 0 ? (long)c * k + 1 : 0) * tail[a];
            rank += (long)(l - (c > 0 ? 1 : 0)) * mprod;
            mprod *= c > 0 ? k : k + 1;
        }
    return pre + rank;
}

// Candidate column points of a row on one axis, in DOF order of their owner
// cells: up to three segments (owner cell, first lattice point, count).  All
// array indices are constants, so the segments stay in registers.
__device__ inline void axis_segments(int p, int k, int n, AxisSegs& s)
{
    int lo, hi;
    if (p % k == 0) { lo = p / k - 1; hi = p / k; } else { lo = hi = p / k; }
    if (lo < 0) lo = 0;
    if (hi > n - 1) hi = n - 1;
    s.clo = lo; s.chi = hi;
    if (lo > 0) {
        s.owner[0] = lo - 1; s.first[0] = k * lo;     s.cnt[0] = 1;
        s.owner[1] = lo;     s.first[1] = k * lo + 1; s.cnt[1] = k;
        s.owner[2] = hi;     s.first[2] = k * hi + 1; s.cnt[2] = k;
        s.nseg = hi > lo ? 3 : 2;
    } else {
        s.owner[0] = 0;  s.first[0] = 0;          s.cnt[0] = k + 1;
        s.owner[1] = hi; s.first[1] = k * hi + 1; s.cnt[1] = k;
        s.owner[2] = 0;  s.first[2] = 0;          s.cnt[2] = 0;
        s.nseg = hi > lo ? 2 : 1;
    }
    s.W = k * (hi - lo + 1) + 1;
}

// x / c for a segment length c in {1, k, k + 1} (k <= 4) and 0 <= x < 8192:
// a multiply by ceil(2^16 / c) and a shift is exact there (checked exhaustively
// for c <= 8), and with a constant k the three multipliers are constants
__device__ inline int div_seg(int x, int c, int k)
{
    const unsigned m = c == 1 ? 0x10001u : (c == k ? 0x10000u / (unsigned)k + 1u : 0x10000u / (unsigned)(k + 1) + 1u);
    return (int)(((unsigned)x * m) >> 16);
}

// decode candidate index t of a row into its column lattice point q
template <int DD, int KK>
__device__ inline void decode_candidate(const GridDev& g, const AxisSegs* S, int t, int* q)
{
    const int d = DK<DD, KK>::d(g);
    int tailW[MAXD];
    int tw = 1;
#pragma unroll
    for (int a = MAXD - 1; a >= 0; a--)
        if (a < d) { tailW[a] = tw; tw *= S[a].W; }
    int rem = t, M = 1, cs[MAXD], fs[MAXD];
#pragma unroll
    for (int a = 0; a < MAXD; a++)
        if (a < d) {
            // segment of axis a: the first whose block (M * cnt * tailW) still contains rem
            int c = S[a].cnt[0], f = S[a].first[0];
            const int b0 = M * S[a].cnt[0] * tailW[a];
            if (S[a].nseg > 1 && rem >= b0) {
                rem -= b0;
                c = S[a].cnt[1];
                f = S[a].first[1];
                const int b1 = M * S[a].cnt[1] * tailW[a];
                if (S[a].nseg > 2 && rem >= b1) {
                    rem -= b1;
                    c = S[a].cnt[2];
                    f = S[a].first[2];
                }
            }
            cs[a] = c;
            fs[a] = f;
            M *= c;
        }
    // rem indexes inside the owner box, axis 0 fastest (mixed radix cs[0], cs[1], ...)
    const int k = DK<DD, KK>::k(g);
#pragma unroll
    for (int a = 0; a < MAXD; a++)
        if (a < d) {
            const int qd = div_seg(rem, cs[a], k);
            q[a] = fs[a] + (rem - qd * cs[a]);
            rem = qd;
        }
}

// value of entry (row node p, column node q) summed over the cells containing
// both, in lexicographic cell order (last axis fastest): sum = integral + sum
// for every cell whose integral is not exactly zero
// value of entry (row node p, column node q) summed over the cells containing
// both, in lexicographic cell order (last axis fastest): sum = integral + sum
// for every cell whose integral is not exactly zero
template <int DD, int KK>
__device__ inline double entry_value(const RowParams& P, const double* __restrict__ E, const int* const* cls,
                                    const int* p, const int* q, bool& present)
{
#pragma clang fp contract(off)
    const int d = DK<DD, KK>::d(P.g), k = DK<DD, KK>::k(P.g);
    // cells containing both p_a and q_a: lo[a] and, if cnt[a] == 2, lo[a] + 1
    int lo[MAXD] = {0, 0, 0}, cnt[MAXD] = {1, 1, 1};
#pragma unroll
    for (int a = 0; a < MAXD; a++)
        if (a < d) {
            int plo, phi, qlo, qhi;
            if (p[a] % k == 0) { plo = p[a] / k - 1; phi = p[a] / k; } else { plo = phi = p[a] / k; }
            if (q[a] % k == 0) { qlo = q[a] / k - 1; qhi = q[a] / k; } else { qlo = qhi = q[a] / k; }
            const int l0 = max(max(plo, qlo), 0), h0 = min(min(phi, qhi), P.g.n[a] - 1);
            lo[a] = l0;
            cnt[a] = h0 - l0 + 1;
        }
    double sum = 0;
    present = false;
    const long nent = (long)P.nloc * P.nloc;
    // lexicographic cell order (last axis fastest); every axis has at most 2 cells
#pragma unroll
    for (int i0 = 0; i0 < 2; i0++)
#pragma unroll
        for (int i1 = 0; i1 < 2; i1++)
#pragma unroll
            for (int i2 = 0; i2 < 2; i2++) {
                const int ii[MAXD] = {i0, i1, i2};
                bool in = true;
#pragma unroll
                for (int a = 0; a < MAXD; a++) in = in && ii[a] < cnt[a];
                if (!in) continue;
                // element-table index < 2^30 and cell index < 2^31 (checked on the host): 32-bit
                unsigned tup = 0, cell = 0;
                int iv = 0, iu = 0, mul = 1;
#pragma unroll
                for (int a = 0; a < MAXD; a++)
                    if (a < d) {
                        const int c = lo[a] + ii[a];
                        tup = tup * (unsigned)P.ncls[a] + (P.ncls[a] == 1 ? 0u : (unsigned)cls[a][c]);
                        cell = cell * (unsigned)P.g.n[a] + (unsigned)c;
                        iv += (p[a] - k * c) * mul;
                        iu += (q[a] - k * c) * mul;
                        mul *= k + 1;
                    }
                double integral = 0;
                for (int i = 0; i < P.nint; i++) {
                    const double e = E[((unsigned)i * (unsigned)P.ntuples + tup) * (unsigned)nent +
                                       (unsigned)(iv * P.nloc + iu)];
                    integral += P.coeff ? P.coeff[cell] * e : e;
                }
                if (integral != 0) {
                    sum = integral + sum;
                    present = true;
                }
            }
    return sum;
}

// G lanes per row; FILL = false counts, FILL = true writes cols/vals
// (FILL also checks that each row received exactly the entries its row pointer
// reserved and raises *bad otherwise)
template <int DD, int KK, int G, bool FILL, bool LDS>
__global__ __launch_bounds__(256) void k_assemble_rows(RowParams P, long nrows, int* __restrict__ counts,
                                                      const int* __restrict__ rowptr, int* __restrict__ cols,
                                                      double* __restrict__ vals, int* __restrict__ bad)
{
#pragma clang fp contract(off)
    const int d = DK<DD, KK>::d(P.g), k = DK<DD, KK>::k(P.g);
    // element tables and cell classes: staged in LDS when they fit (uniform and
    // few-class grids), read from global memory otherwise
    extern __shared__ double lds_tab[];
    const double* E = P.E;
    const int* cls[MAXD] = {P.cls[0], P.cls[1], P.cls[2]};
    if constexpr (LDS) {
        const long esize = (long)P.nint * P.ntuples * P.nloc * P.nloc;
        for (long i = threadIdx.x; i < esize; i += blockDim.x) lds_tab[i] = P.E[i];
        int* lc = reinterpret_cast<int*>(lds_tab + esize);
        int off = 0;
#pragma unroll
        for (int a = 0; a < MAXD; a++)
            if (a < d) {
                for (int i = threadIdx.x; i < P.g.n[a]; i += blockDim.x) lc[off + i] = P.cls[a][i];
                off += P.g.n[a];
            }
        __syncthreads();
        E = lds_tab;
        off = 0;
#pragma unroll
        for (int a = 0; a < MAXD; a++)
            if (a < d) { cls[a] = lc + off; off += P.g.n[a]; }
    }
    const int lane = threadIdx.x % G;
    const int wlane = threadIdx.x & 63;
    const int gbase = wlane - lane;   // first lane of the group inside the wavefront
    const long groups_total = (long)gridDim.x * (blockDim.x / G);
    const long gi = (blockIdx.x * (long)blockDim.x + threadIdx.x) / G;
    const long wave_first = gi - wlane / G;   // first group of this wavefront
    long tails[MAXD] = {1, 1, 1};
    lattice_tails<DD, KK>(P.g, tails);
    for (long it = 0; wave_first + it < nrows; it += groups_total) {
        // with one row per wavefront the row is uniform: its setup (lattice point,
        // segments, cells, element-table offsets) then lives in scalar registers
        const long row = G == 64 ? (long)__builtin_amdgcn_readfirstlane((int)(gi + it)) : gi + it;
        // all lanes of a wavefront iterate together so that ballots stay uniform
        const bool active = row < nrows;
        AxisSegs S[MAXD];
        int p[MAXD] = {0, 0, 0};
        int total = 0;
        if (active) {
            lat_of_dof_t<DD, KK>(P.g, row, p);
            total = 1;
#pragma unroll
            for (int a = 0; a < MAXD; a++)
                if (a < d) { axis_segments(p[a], k, P.g.n[a], S[a]); total *= S[a].W; }
        }
        // uniform trip count across the wavefront
        int my_chunks = (total + G - 1) / G;
        int chunks = my_chunks;
        for (int o = 32; o > 0; o >>= 1) chunks = max(chunks, __shfl_xor(chunks, o, 64));
        int written = 0;
        const int base = (FILL && active) ? rowptr[row] : 0;
        for (int ch = 0; ch < chunks; ch++) {
            const int t = ch * G + lane;
            bool present = false;
            double v = 0;
            int colv = 0;
            if (active && t < total) {
                int q[MAXD] = {0, 0, 0};
                decode_candidate<DD, KK>(P.g, S, t, q);
                v = entry_value<DD, KK>(P, E, cls, p, q, present);
                if (FILL || row == 0) colv = (int)dof_of_lat_tails<DD, KK>(P.g, tails, q);
                if (row == 0 && colv == 0) present = true;   // dummy (0,0) entry
            }
            const unsigned long long bal = __ballot(present);
            const unsigned long long gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << gbase);
            const unsigned long long mine = bal & gmask;
            if (FILL && present) {
                const unsigned long long below = mine & ((1ull << wlane) - 1ull);
                const int pos = base + written + __popcll(below);
                cols[pos] = colv;
                vals[pos] = v;
            }
            written += __popcll(mine);
        }
        if (!FILL && active && lane == 0) counts[row] = written;
        if (FILL && active && lane == 0 && written != rowptr[row + 1] - base) atomicOr(bad, 1);
    }
}

// ---- cell-major row assembly (d <= 3, k <= 2) ----
// Per axis a row's node p lies in one or two cells lo..hi, and its candidate columns are the
// lattice points k*lo .. k*hi + k, at offsets o = q - k*lo.  The order of the candidates (the
// DOF order of their owner cells, which k_assemble_rows decodes per candidate) depends only on
// the row's type per axis: (lo > 0, hi > lo).  The host tabulates, for each combination of
// types, candidate t -> offsets (fwd, 4 bits per axis) and offsets -> t (inv).  A row then runs
// over its own cells in lexicographic order (last axis fastest); lane l takes local node l of
// the cell and adds the cell's integral to its candidate's sum in LDS.  The cells containing
// both nodes of an entry are exactly the row's cells that contain the column node, visited in
// the same lexicographic order as entry_value, so every entry is summed with the same
// operations in the same order (bit-identical).
constexpr int CA_MAXW = 125;   // (2k + 1)^d candidates for d = 3, k = 2

template <int DD, int KK>
struct CellAsm {
    static constexpr int K1 = KK + 1, W1 = 2 * KK + 1;
    static constexpr int NLOC = DD == 3 ? K1 * K1 * K1 : DD == 2 ? K1 * K1 : K1;
    static constexpr int WMAX = DD == 3 ? W1 * W1 * W1 : DD == 2 ? W1 * W1 : W1;
    static constexpr int NCOMBO = 1 << (2 * DD);
};

// What a row needs of one axis a at its lattice coordinate p (tabulated on the host by
// cell_asm_axes, 32 bytes, so one scalar load brings it): the first cell lo containing p; info =
// nc | tau << 2 | class << 4 with nc the cells containing p (1 or 2), tau the type bits
// 2 [lo > 0] + [hi > lo], class 0 a vertex between two cells, 1 a node inside a cell, 2 the last
// vertex (when lo >= 2), 3 (lo < 2); and per cell j < nc its part of the element-table offset
// (cell class * stride of the class-tuple index * (k+1)^2d + local index of p * (k+1)^d) and of
// the cell index.
struct __align__(32) AxEntry {
    int lo, info;
    unsigned L[2], C[2];
    int pad[2];
};

struct AxOffsets {
    int o[MAXD];
};

// One wavefront per segment of seg_len nodes of a lattice line (every axis but the last fixed;
// nseg segments per line; line_order lists the lines grouped by the classes of their fixed axes'
// cells, so that the lines in flight together read few element tables, which then stay in L2);
// G lanes per row, 64 / G consecutive nodes at a time.  The fixed
// axes' parts of the row's cells are summed once per line.  Per row the loads go out first
// (row bounds, candidate offsets, and per cell of the row the candidate index of the lane's
// node, the coefficient and the element-table entry); the sums then run over the cells in
// order.  A candidate's sum starts at -0.0: x + (-0.0) = x for every x, and a sum of nonzero
// terms is never -0.0 in round-to-nearest, so "present" is "the sum is not -0.0".
// Columns: on an inner row (every axis' cells have index >= 2, so every candidate's owner cell
// has index >= 1) col - row depends only on the row's axis classes and the candidate (the DOF
// is sum_a k^a (q_a tail_a + (tail_a - 1)(1 - l_a)) there, l_a the node's index in its owner
// cell), tabulated on the host in delta; other rows evaluate the closed form per candidate.
template <int DD, int KK, int G, bool FILL, bool LDS, bool COEF>
__global__ __launch_bounds__(256) void k_asm_cells(RowParams P, long nunits, int nseg, int seg_len,
                                                  int* __restrict__ counts,
                                                  const int* __restrict__ rowptr, int* __restrict__ cols,
                                                  double* __restrict__ vals, int* __restrict__ bad,
                                                  const unsigned short* __restrict__ fwd,
                                                  const unsigned char* __restrict__ inv,
                                                  const int* __restrict__ delta, const AxEntry* __restrict__ axtab,
                                                  AxOffsets aoff, const int* __restrict__ line_order, int inner_ok)
{
#pragma clang fp contract(off)
    using CA = CellAsm<DD, KK>;
    constexpr int NLOC = CA::NLOC, WMAX = CA::WMAX, W1 = CA::W1, K1 = CA::K1;
    constexpr int R = 64 / G;                         // rows per wavefront at a time
    constexpr int CHUNKS = (WMAX + G - 1) / G;
    constexpr int NH = 1 << DD;                       // at most two cells per axis contain a node
    constexpr int NF = NH / 2;                        // cell slots of the fixed axes
    constexpr int LAST = DD - 1;
    constexpr int WLAST = DD == 3 ? W1 * W1 : DD == 2 ? W1 : 1;   // W1^LAST
    constexpr int C3LAST = DD == 3 ? 9 : DD == 2 ? 3 : 1;         // 3^LAST
    // per row group: the candidates' sums and one slot (STRIDE - 1) for the lanes that hold no
    // node of a cell, so that every lane can take part in the sums' stores
    constexpr int STRIDE = (WMAX + 1 + G - 1) / G * G;
    static_assert(NLOC <= G && G <= 64 && WMAX <= CA_MAXW && CHUNKS * G <= STRIDE, "one lane per local node");
    __shared__ double acc_all[2 * (256 / G) * STRIDE];   // two rows in flight per row group
    extern __shared__ double lds_tab[];
    const double* E = P.E;
    if constexpr (LDS) {
        const long esize = (long)P.nint * P.ntuples * NLOC * NLOC;
        for (long i = threadIdx.x; i < esize; i += blockDim.x) lds_tab[i] = P.E[i];
        E = lds_tab;
        __syncthreads();
    }
    const int lane = threadIdx.x % G;
    const int sub = (threadIdx.x & 63) / G;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double* acc = acc_all + (threadIdx.x / G) * STRIDE;
    const long unit = blockIdx.x * 4L + wave;
    if (unit >= nunits) return;   // wave-uniform
    const long li = unit / nseg;
    const long line = line_order[li];
    const int seg0 = (int)(unit - li * nseg) * seg_len;
    const int wlane = threadIdx.x & 63;
    const int gbase = wlane - lane;
    long tails[MAXD] = {1, 1, 1};
    lattice_tails<DD, KK>(P.g, tails);
    int p[MAXD] = {0, 0, 0};
    AxEntry ax[MAXD];
    {
        long r = line;
#pragma unroll
        for (int a = LAST - 1; a >= 0; a--) {
            const int Na = KK * P.g.n[a] + 1;
            p[a] = (int)(r % Na);
            r /= Na;
        }
#pragma unroll
        for (int a = 0; a < LAST; a++) ax[a] = axtab[aoff.o[a] + p[a]];
    }
    // the fixed axes' part of each cell slot hf (their cell bits, first axis most significant);
    // vmask: bits 2 hf and 2 hf + 1 set for the slots whose cells exist
    unsigned LF[NF], CF[NF];
    int HF[NF];
    int vmask = 0;
#pragma unroll
    for (int hf = 0; hf < NF; hf++) {
        unsigned ll = 0, cc = 0;
        int ho = 0, wm = 1;
        bool v = true;
#pragma unroll
        for (int a = 0; a < LAST; a++) {
            const int j = (hf >> (LAST - 1 - a)) & 1;
            v = v && j < (ax[a].info & 3);
            ll += ax[a].L[j];
            cc += ax[a].C[j];
            ho += j * KK * wm;
            wm *= W1;
        }
        LF[hf] = ll; CF[hf] = cc; HF[hf] = ho;
        if (v) vmask |= 3 << (2 * hf);
    }
    int combo_f = 0, total_f = 1, class_f = 0, cw = 1;
    bool inner_f = inner_ok != 0;
#pragma unroll
    for (int a = 0; a < LAST; a++) {
        combo_f |= ((ax[a].info >> 2) & 3) << (2 * a);
        total_f *= KK * (ax[a].info & 3) + 1;
        const int cl = ax[a].info >> 4;
        inner_f = inner_f && cl < 3;
        class_f += cl * cw;
        cw *= 3;
    }
    // the row's DOF along the line: dof(p) = dof(p with p_last = 0) + mprod * p_last, mprod the
    // product over the fixed axes of the owner-cell extents (k, or k + 1 in cell 0)
    p[LAST] = 0;
    const long rbase = dof_of_lat_tails<DD, KK>(P.g, tails, p);
    int rstep = 1;
#pragma unroll
    for (int a = 0; a < LAST; a++) rstep *= (p[a] == 0 ? 0 : (p[a] - 1) / KK) > 0 ? KK : KK + 1;
    // this lane's local node (axis 0 fastest) as an offset in the candidate box
    int lofs = 0;
    {
        int r = lane, wm = 1;
#pragma unroll
        for (int a = 0; a < DD; a++) { lofs += (r % K1) * wm; r /= K1; wm *= W1; }
    }
    const bool node_lane = lane < NLOC;
    const unsigned ioff = (unsigned)P.ntuples * (unsigned)(NLOC * NLOC);   // between the integrals' tables
    const int NL = min(KK * P.g.n[LAST] + 1, seg0 + seg_len);
    const int iters = (NL - seg0 + R - 1) / R;
    // A row's inputs.  Two rows of the line are in flight at a time: both rows' loads go out
    // before either is used, their sums interleave, then both rows are written.
    struct RowIn {
        bool active, inner;
        int M, combo, total, lo, base, lim;
        long row;
        unsigned L[2];
        unsigned fo[CHUNKS];
        int dl[CHUNKS];
        int ht[NH];
        double ev[NH], cf[NH];
    };
    auto fetch = [&](int it, RowIn& r) {
        const int pl = seg0 + it * R + sub;
        r.active = pl < NL;
        const int pla = r.active ? pl : seg0;
        const AxEntry xl = axtab[aoff.o[LAST] + pla];
        const int ncl = xl.info & 3, cll = xl.info >> 4;
        r.combo = combo_f | (((xl.info >> 2) & 3) << (2 * LAST));
        r.total = total_f * (KK * ncl + 1);
        r.row = rbase + (long)rstep * pla;
        r.inner = inner_f && cll < 3;
        r.lo = xl.lo;
        r.L[0] = xl.L[0];
        r.L[1] = xl.L[1];
        const int dcls = r.inner ? class_f + cll * C3LAST : 0;
        r.M = r.active ? (vmask & (ncl == 2 ? 0xff : 0x55)) : 0;   // the row's cell slots
        r.base = r.lim = 0;
        if (FILL) {
            r.base = rowptr[r.row];
            r.lim = rowptr[r.row + 1];
            if (!r.active) r.base = r.lim = 0;
        }
#pragma unroll
        for (int ch = 0; ch < CHUNKS; ch++) {
            r.fo[ch] = fwd[r.combo * WMAX + ch * G + lane];
            r.dl[ch] = delta[dcls * WMAX + ch * G + lane];
        }
        // the row's cells: slot h = (fixed-axes slot hf, last-axis bit j), lexicographic order.
        // With one row per wavefront the slot mask is uniform: skipped slots cost a branch.
        // All lanes load (lanes past the cell's nodes read padding) and all take part in the
        // sums; those lanes add into the spare slot.
#pragma unroll
        for (int h = 0; h < NH; h++) {
            const int hf = h >> 1, j = h & 1;
            if ((r.M >> h) & 1) {
                r.cf[h] = COEF ? P.coeff[CF[hf] + xl.C[j]] : 1.0;
                r.ht[h] = inv[r.combo * WMAX + HF[hf] + j * KK * WLAST + lofs];
                // element-table index < 2^30 (checked on the host): 32-bit
                r.ev[h] = E[LF[hf] + xl.L[j] + (unsigned)lane];
            }
        }
    };
    // the sums of one row over its cells in their order into acc
    auto sum_row = [&](const RowIn& r, double* acc, int h) {
        const int hf = h >> 1, j = h & 1;
        if ((r.M >> h) & 1) {
            double x = 0;
            x += COEF ? r.cf[h] * r.ev[h] : r.ev[h];
            for (int i = 1; i < P.nint; i++) {   // further integrals: E of integral i at i * ioff
                const double e = E[(unsigned)i * ioff + LF[hf] + r.L[j] + (unsigned)lane];
                x += COEF ? r.cf[h] * e : e;
            }
            const int t = node_lane ? r.ht[h] : STRIDE - 1;
            const double old = acc[t];
            acc[t] = x != 0 ? x + old : old;
        }
    };
    // candidates in column order: ballot-compacted writes (G lanes per row).  Per-lane
    // conditions become selects; the one branch per chunk is around the stores.
    auto write_row = [&](const RowIn& r, const double* acc) {
        int written = 0;
#pragma unroll
        for (int ch = 0; ch < CHUNKS; ch++) {
            const int t = ch * G + lane;
            double v = acc[t];
            bool present = r.active && t < r.total &&
                           __double_as_longlong(v) != (long long)0x8000000000000000ull;
            int colv = 0;
            if (FILL || r.row == 0) {
                if (r.inner) {
                    colv = (int)r.row + r.dl[ch];
                } else {
                    int q[MAXD] = {0, 0, 0};
                    q[LAST] = KK * r.lo + (int)((r.fo[ch] >> (4 * LAST)) & 15u);
#pragma unroll
                    for (int a = 0; a < LAST; a++) q[a] = KK * ax[a].lo + (int)((r.fo[ch] >> (4 * a)) & 15u);
                    colv = (int)dof_of_lat_tails<DD, KK>(P.g, tails, q);
                }
            }
            if (r.row == 0) {   // the dummy (0,0) entry of a new matrix
                const bool dummy = r.active && t < r.total && colv == 0 && !present;
                present = present || dummy;
                v = dummy ? 0.0 : v;
            }
            const unsigned long long bal = __ballot(present);
            const unsigned long long gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << gbase);
            const unsigned long long mine = bal & gmask;
            if (FILL) {
                const unsigned long long below = mine & ((1ull << wlane) - 1ull);
                const int pos = r.base + written + __popcll(below);
                if (present && pos < r.lim) {   // a row never writes past its reserved range
                    cols[pos] = colv;
                    vals[pos] = v;
                }
            }
            written += __popcll(mine);
        }
        if (!FILL && r.active && lane == 0) counts[r.row] = written;
        if (FILL && r.active && lane == 0 && written != r.lim - r.base) atomicOr(bad, 1);
    };
    double* acc1 = acc_all + (256 / G) * STRIDE + (threadIdx.x / G) * STRIDE;
    for (int it = 0; it < iters; it += 2) {
        const bool two = it + 1 < iters;
        RowIn ra, rb;
        fetch(it, ra);
        fetch(two ? it + 1 : it, rb);
        if (!two) rb.M = 0, rb.active = false;
#pragma unroll
        for (int k2 = 0; k2 < STRIDE / G; k2++) {
            acc[k2 * G + lane] = -0.0;
            acc1[k2 * G + lane] = -0.0;
        }
#pragma unroll
        for (int h = 0; h < NH; h++) {
            sum_row(ra, acc, h);
            sum_row(rb, acc1, h);
        }
        write_row(ra, acc);
        write_row(rb, acc1);
    }
}

// the per-axis node table of k_asm_cells (AxEntry) for every axis, axes one after another
static void cell_asm_axes(const Grid& g, const std::vector<int>* cls, const int* ncls, std::vector<AxEntry>& tab,
                          AxOffsets& off)
{
    const int d = g.d, k = g.k;
    int nloc = 1;
    for (int a = 0; a < d; a++) nloc *= k + 1;
    const unsigned nent = (unsigned)(nloc * nloc);
    unsigned ctup[MAXD], civ[MAXD];
    {
        unsigned st = 1;
        for (int a = d - 1; a >= 0; a--) { ctup[a] = st; st *= (unsigned)ncls[a]; }
        unsigned m = 1;
        for (int a = 0; a < d; a++) { civ[a] = m; m *= (unsigned)(k + 1); }
    }
    unsigned ccell[MAXD];
    {
        unsigned sc = 1;
        for (int a = d - 1; a >= 0; a--) { ccell[a] = sc; sc *= (unsigned)g.n[a]; }
    }
    tab.clear();
    for (int a = 0; a < MAXD; a++) off.o[a] = 0;
    for (int a = 0; a < d; a++) {
        off.o[a] = (int)tab.size();
        const int N = k * g.n[a] + 1;
        for (int p = 0; p < N; p++) {
            int l, h;
            if (p % k == 0) { l = p / k - 1; h = p / k; } else { l = h = p / k; }
            l = std::max(l, 0);
            h = std::min(h, g.n[a] - 1);
            const int nc = h - l + 1;
            const int tau = (l > 0 ? 2 : 0) | (h > l ? 1 : 0);
            const int cl = l < 2 ? 3 : (tau == 3 ? 0 : (p - k * l < k ? 1 : 2));
            AxEntry e;
            std::memset(&e, 0, sizeof(e));
            e.lo = l;
            e.info = nc | tau << 2 | cl << 4;
            for (int j = 0; j < 2; j++) {
                const int c = std::min(l + j, h);   // j = 1 with one cell: a valid index, never used
                const unsigned cv = ncls[a] == 1 ? 0u : (unsigned)cls[a][c];
                e.L[j] = cv * ctup[a] * nent + (unsigned)(p - k * c) * civ[a] * (unsigned)nloc;
                e.C[j] = (unsigned)c * ccell[a];
            }
            tab.push_back(e);
        }
    }
}

// smallest |E| over the element tables (bit patterns of non-negative doubles
// order like the values, so an integer atomicMin is exact and deterministic)
__global__ void k_absmin(long n, const double* __restrict__ e, unsigned long long* __restrict__ out)
{
    unsigned long long m = ~0ull;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(e[i]));
        m = b < m ? b : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(m, o, 64);
        m = t < m ? t : m;
    }
    // one atomic per workgroup (atomics on one address serialise)
    __shared__ unsigned long long wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = wm[w] < m ? wm[w] : m;
        atomicMin(out, m);
    }
}

// row lengths when every structural candidate is present: prod_a W_a
template <int DD, int KK>
__global__ __launch_bounds__(256) void k_count_structural(GridDev g, long nrows, int* __restrict__ counts)
{
    const int d = DK<DD, KK>::d(g), k = DK<DD, KK>::k(g);
    for (long row = blockIdx.x * 256L + threadIdx.x; row < nrows; row += (long)gridDim.x * 256) {
        int p[MAXD] = {0, 0, 0};
        lat_of_dof_t<DD, KK>(g, row, p);
        int tot = 1;
#pragma unroll
        for (int a = 0; a < MAXD; a++)
            if (a < d) {
                AxisSegs S;
                axis_segments(p[a], k, g.n[a], S);
                tot *= S.W;
            }
        counts[row] = tot;
    }
}

// --------------------------- scan (int64 totals) ---------------------------
constexpr int SCAN_BS = 256;
constexpr int SCAN_PER = 8;
constexpr int SCAN_TILE = SCAN_BS * SCAN_PER;

__device__ inline long block_excl_scan(long v, long* sh, long& total)
{
    // inclusive scan in wavefront, then across wavefronts
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        long acc = 0;
        for (int i = 0; i < SCAN_BS / 64; i++) { const long t = sh[i]; sh[i] = acc; acc += t; }
        sh[SCAN_BS / 64] = acc;
    }
    __syncthreads();
    total = sh[SCAN_BS / 64];
    const long r = x - v + sh[w];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_tiles(const int* __restrict__ in, long n, long* __restrict__ out,
                                                        long* __restrict__ tile_sums)
{
    __shared__ long sh[SCAN_BS / 64 + 1];
    const long base = (long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    long loc[SCAN_PER];
    long s = 0;
    for (int i = 0; i < SCAN_PER; i++) { loc[i] = (base + i < n) ? in[base + i] : 0; s += loc[i]; }
    long total;
    long off = block_excl_scan(s, sh, total);
    for (int i = 0; i < SCAN_PER; i++) { if (base + i < n) out[base + i] = off; off += loc[i]; }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_sums(long* __restrict__ sums, long m, long* __restrict__ grand)
{
    __shared__ long sh[SCAN_BS / 64 + 1];
    long carry = 0;
    for (long b = 0; b < m; b += SCAN_BS) {
        const long i = b + threadIdx.x;
        const long v = i < m ? sums[i] : 0;
        long total;
        const long e = block_excl_scan(v, sh, total);
        if (i < m) sums[i] = e + carry;
        carry += total;
    }
    if (threadIdx.x == 0) *grand = carry;
}

// rowptr[i] = tile-local prefix + tile offset; rowptr[n] = the grand total (only read where it
// is below 2^31)
__global__ void k_scan_finish(const long* __restrict__ tmp, const long* __restrict__ sums, const long* __restrict__ grand,
                              long n, int* __restrict__ rowptr)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i <= n; i += (long)gridDim.x * blockDim.x)
        rowptr[i] = i < n ? (int)(tmp[i] + sums[i / SCAN_TILE]) : (int)*grand;
}

size_t scan_work_bytes(long n)
{
    const long tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    return sizeof(long) * ((n > 0 ? n : 1) + tiles + 2);
}

long scan_counts(jfem_ctx* c, const int* counts, long n, int* rowptr, long* work, long known_total, hipStream_t st)
{
    if (!st) st = c->stream;
    const long tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    long* tmp = work;
    long* sums = work + (n > 0 ? n : 1);
    if (n > 0) k_scan_tiles<<<tiles, SCAN_BS, 0, st>>>(counts, n, tmp, sums);
    k_scan_sums<<<1, SCAN_BS, 0, st>>>(sums, tiles, sums + tiles + 1);
    long total = known_total;
    if (known_total < 0) {
        JFEM_HIP(hipMemcpyAsync(c->pin, sums + tiles + 1, sizeof(long), hipMemcpyDeviceToHost, st));
        JFEM_HIP(hipStreamSynchronize(st));
        total = c->pin[0];
    }
    JFEM_CHECK(total >= 0 && total < (1l << 31) - 1, JFEM_ERR_UNSUPPORTED,
               "matrix has 2^31 or more entries (int32 CSR indices)");
    k_scan_finish<<<grid_for(n + 1, 256, 8192), 256, 0, st>>>(tmp, sums, sums + tiles + 1, n, rowptr);
    JFEM_HIP(hipGetLastError());
    return total;
}

// rowptr[0..n] from counts[0..n-1]; returns total (host)
long scan_counts(jfem_ctx* c, const int* counts, long n, int* rowptr)
{
    long* work = nullptr;
    JFEM_HIP(hipMalloc(&work, scan_work_bytes(n)));
    long total = -1;
    try {
        total = scan_counts(c, counts, n, rowptr, work, -1);
    } catch (...) {
        (void)hipFree(work);   // waits for the queued scan kernels
        throw;
    }
    (void)hipFree(work);
    return total;
}

// ------------------------- space helper kernels ----------------------------

__global__ void k_cell_dofs(GridDev g, long ncells, int nloc, int* __restrict__ out)
{
    for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < ncells * nloc; id += (long)gridDim.x * blockDim.x) {
        const long cell = id / nloc;
        int li = (int)(id % nloc);
        int c[MAXD], p[MAXD];
        long r = cell;
        for (int a = g.d - 1; a >= 0; a--) { c[a] = (int)(r % g.n[a]); r /= g.n[a]; }
        for (int a = 0; a < g.d; a++) { p[a] = c[a] * g.k + li % (g.k + 1); li /= g.k + 1; }
        out[id] = (int)dof_of_lattice(g, p);
    }
}

__global__ void k_boundary_mask(GridDev g, long n, unsigned char* __restrict__ mask)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int p[MAXD];
        lattice_of_dof(g, i, p);
        bool b = false;
        for (int a = 0; a < g.d; a++) b = b || p[a] == 0 || p[a] == g.k * g.n[a];
        mask[i] = b ? 1 : 0;
    }
}

struct Coord1D { double s[MAXD]; double e[MAXD]; };

// node coordinates: (1.0/k * l) * len + start of the owner cell (LagrangeBasisFunction1D ctor :19)
__global__ void k_node_coords(GridDev g, long n, const double* __restrict__ s0, const double* __restrict__ s1,
                              const double* __restrict__ s2, const double* __restrict__ e0,
                              const double* __restrict__ e1, const double* __restrict__ e2, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const double* S[MAXD] = {s0, s1, s2};
    const double* Ee[MAXD] = {e0, e1, e2};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int p[MAXD];
        lattice_of_dof(g, i, p);
        for (int a = 0; a < g.d; a++) {
            const int c = p[a] == 0 ? 0 : (p[a] - 1) / g.k;
            const int l = p[a] - c * g.k;
            out[i * g.d + a] = 1.0 / g.k * l * (Ee[a][c] - S[a][c]) + S[a][c];
        }
    }
}

// ScalarFESpaceFunction.value (M/basic/ScalarFESpaceFunction.java:40-45) at
// a point: sum over the shape functions of coefficient * value.  A shape
// function's value is its product of 1D Lagrange values in the first of its
// cells that contains the point (ScalarShapeFunction.value, Cell1D.isInCell
// with 1e-8 slack), 0 if none does; only the (k+1)^d functions of the point's
// cell are nonzero.  Per axis the cell is the one the point's coordinate falls
// in (the lower one on a shared face, where continuity makes both agree up to
// rounding); a point outside the domain evaluates to 0.
__global__ void k_eval_fe(GridDev g, const double* __restrict__ s0, const double* __restrict__ s1,
                          const double* __restrict__ s2, const double* __restrict__ e0, const double* __restrict__ e1,
                          const double* __restrict__ e2, const double* __restrict__ coeff, long npts,
                          const double* __restrict__ pts, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const double* S[MAXD] = {s0, s1, s2};
    const double* Ee[MAXD] = {e0, e1, e2};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npts; i += (long)gridDim.x * blockDim.x) {
        int c[MAXD] = {0, 0, 0};
        double t[MAXD] = {0, 0, 0};
        bool inside = true;
        for (int a = 0; a < g.d; a++) {
            const double x = pts[i * g.d + a];
            const double w = (Ee[a][g.n[a] - 1] - S[a][0]) / g.n[a];
            int ca = (int)floor((x - S[a][0]) / w);
            ca = ca < 0 ? 0 : (ca > g.n[a] - 1 ? g.n[a] - 1 : ca);
            // the rounded division can land one cell off near a face
            if (!in_cell(S[a][ca], Ee[a][ca], x) && ca > 0 && in_cell(S[a][ca - 1], Ee[a][ca - 1], x)) ca--;
            if (!in_cell(S[a][ca], Ee[a][ca], x) && ca < g.n[a] - 1 && in_cell(S[a][ca + 1], Ee[a][ca + 1], x)) ca++;
            if (!in_cell(S[a][ca], Ee[a][ca], x)) inside = false;
            c[a] = ca;
            t[a] = (x - S[a][ca]) / (Ee[a][ca] - S[a][ca]);
        }
        double v = 0;
        if (inside) {
            int nloc = 1;
            for (int a = 0; a < g.d; a++) nloc *= g.k + 1;
            for (int li = 0; li < nloc; li++) {
                int p[MAXD] = {0, 0, 0};
                double phi = 1;
                int r = li;
                for (int a = 0; a < g.d; a++) {
                    const int l = r % (g.k + 1);
                    r /= g.k + 1;
                    p[a] = c[a] * g.k + l;
                    phi *= lag_value(t[a], l, g.k);
                }
                v += phi * coeff[dof_of_lattice(g, p)];
            }
        }
        out[i] = v;
    }
}

// v1 - v2 per point
__global__ void k_diff(long n, const double* a, const double* __restrict__ b, double* out)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = a[i] - b[i];
}

static void cells_1d(double start, double end, int n, std::vector<double>& s, std::vector<double>& e);

// device values of an FE function at npts device points
void space_evaluate_dev(jfem_space* s, const double* coeff, long npts, const double* pts_dev, double* out_dev)
{
    const Grid& g = s->grid;
    DevBuf<double> se[2 * MAXD];
    for (int a = 0; a < g.d; a++) {
        std::vector<double> s1, e1;
        cells_1d(g.start[a], g.end[a], g.n[a], s1, e1);
        se[a].alloc(g.n[a]);
        se[MAXD + a].alloc(g.n[a]);
        JFEM_HIP(hipMemcpyAsync(se[a], s1.data(), sizeof(double) * g.n[a], hipMemcpyHostToDevice, s->ctx->stream));
        JFEM_HIP(hipMemcpyAsync(se[MAXD + a], e1.data(), sizeof(double) * g.n[a], hipMemcpyHostToDevice,
                                s->ctx->stream));
        JFEM_HIP(hipStreamSynchronize(s->ctx->stream));   // s1, e1 leave scope
    }
    if (npts > 0)
        k_eval_fe<<<grid_for(npts, 256, 16384), 256, 0, s->ctx->stream>>>(to_dev(g), se[0], se[1], se[2], se[3], se[4],
                                                                         se[5], coeff, npts, pts_dev, out_dev);
    JFEM_HIP(hipGetLastError());
    JFEM_HIP(hipStreamSynchronize(s->ctx->stream));
}

// ------------------------------ RHS (sines) --------------------------------
// TPRightHandSideIntegral VALUE: per cell, ret_v = sum_q (phi_v(x_q) f(x_q)) w_0 .. w_{d-1};
// rhs[dof] += ret_v.  Cells of one parity colour share no node, so 2^d passes
// accumulate without atomics (summation order per DOF: colour order).
struct RhsParams {
    GridDev g;
    int Q, nloc, nterms;
    const int* cls[MAXD];
    int ncls[MAXD];
    const double* sinq[MAXD];   // sin(kw * pi * x_q) per axis: [term][cell][q]
};

// One workgroup of 256 threads takes CPB = 256 / nloc cells of one colour:
// first the source at the cells' quadrature points goes to LDS, then thread
// (cell, iv) sums its basis function's quadrature (the arithmetic of each value
// in the same order as before).  DD, KK, QQ > 0 fix dimension, degree and
// points per axis, so the point loops unroll.
constexpr int RHS_BS = 256;

template <int DD, int KK, int QQ>
__global__ __launch_bounds__(RHS_BS) void k_rhs_color(RhsParams P, const Tab1D* __restrict__ t0,
                                                      const Tab1D* __restrict__ t1, const Tab1D* __restrict__ t2,
                                                      int color, long ncolor_cells, double* __restrict__ rhs)
{
#pragma clang fp contract(off)
    extern __shared__ double fq[];   // [CPB][nq]
    const int d = DD ? DD : P.g.d, k = KK ? KK : P.g.k, Q = QQ ? QQ : P.Q;
    int nq = 1;
    for (int a = 0; a < d; a++) nq *= Q;
    const int nloc = P.nloc, cpb = RHS_BS / nloc;
    int half[MAXD] = {1, 1, 1};
    for (int a = 0; a < d; a++) half[a] = (P.g.n[a] - ((color >> a) & 1) + 1) / 2;
    for (long cb = (long)blockIdx.x * cpb; cb < ncolor_cells; cb += (long)gridDim.x * cpb) {
        __syncthreads();
        // the source at the quadrature points of the block's cells
        for (int f = threadIdx.x; f < cpb * nq; f += RHS_BS) {
            const long cid = cb + f / nq;
            if (cid >= ncolor_cells) continue;
            int c[MAXD] = {0, 0, 0};
            long r = cid;
            for (int a = d - 1; a >= 0; a--) { c[a] = 2 * (int)(r % half[a]) + ((color >> a) & 1); r /= half[a]; }
            int qa[MAXD] = {0, 0, 0};
            int qq = f % nq;
            for (int a = d - 1; a >= 0; a--) { qa[a] = qq % Q; qq /= Q; }
            double sv = 0;
            for (int m = 0; m < P.nterms; m++) {
                double t = 1;   // amplitude folded into the axis-0 table
                for (int a = 0; a < d; a++) t *= P.sinq[a][((long)m * P.g.n[a] + c[a]) * Q + qa[a]];
                sv += t;
            }
            fq[f] = sv;
        }
        __syncthreads();
        const int ci = threadIdx.x / nloc, iv = threadIdx.x % nloc;
        const long cid = cb + ci;
        if (ci >= cpb || cid >= ncolor_cells) continue;
        int c[MAXD] = {0, 0, 0};
        long r = cid;
        for (int a = d - 1; a >= 0; a--) { c[a] = 2 * (int)(r % half[a]) + ((color >> a) & 1); r /= half[a]; }
        const Tab1D* tb[MAXD] = {t0 + P.cls[0][c[0]], d > 1 ? t1 + P.cls[1][c[1]] : t0,
                                 d > 2 ? t2 + P.cls[2][c[2]] : t0};
        int lv[MAXD] = {0, 0, 0}, p[MAXD] = {0, 0, 0};
        int rr = iv;
#pragma unroll
        for (int a = 0; a < MAXD; a++)
            if (a < d) { lv[a] = rr % (k + 1); rr /= k + 1; p[a] = c[a] * k + lv[a]; }
        const double* fc = fq + ci * nq;
        constexpr int QN = QQ && DD ? (DD == 3 ? QQ * QQ * QQ : DD == 2 ? QQ * QQ : QQ) : 1;
        double ret = 0;
#pragma unroll QN
        for (int q = 0; q < (QN > 1 ? QN : nq); q++) {
            int qa[MAXD] = {0, 0, 0};
            int qq = q;
#pragma unroll
            for (int a = MAXD - 1; a >= 0; a--)
                if (a < d) { qa[a] = qq % Q; qq /= Q; }
            double v = 1;
#pragma unroll
            for (int a = 0; a < MAXD; a++)
                if (a < d) v *= tb[a]->val[lv[a]][qa[a]];
            double val = v * fc[q];
#pragma unroll
            for (int a = 0; a < MAXD; a++)
                if (a < d) val *= tb[a]->w[qa[a]];
            ret += val;
        }
        const long dof = dof_of_lattice(P.g, p);
        if (ret != 0) rhs[dof] += ret;
    }
}

// ------------------------------ host driver --------------------------------

struct ClassTables {
    int Q = 0;
    int ncls[MAXD] = {1, 1, 1};
    std::vector<int> cls[MAXD];        // per 1D cell
    std::vector<Tab1D> tabs[MAXD];     // per class
    std::vector<double> s1d[MAXD], e1d[MAXD];
};

// CartesianGrid.createCells1D (M/tensorproduct/geometry/CartesianGrid.java:26-49)
#pragma clang fp contract(off)
static void cells_1d(double start, double end, int n, std::vector<double>& s, std::vector<double>& e)
{
    const double width = (end - start) / n;
    s.resize(n); e.resize(n);
    for (int j = 0; j < n; j++) { s[j] = j * width + start; e[j] = (j + 1) * width + start; }
}
#pragma clang fp contract(on)

static ClassTables build_classes(const Grid& g, int rule)
{
    ClassTables C;
    double qp[8], qw[8];
    C.Q = quad_rule(rule, qp, qw);
    JFEM_CHECK(C.Q > 0, JFEM_ERR_ARG, "unknown quadrature rule");
    JFEM_CHECK(g.k >= 1 && g.k < MAXK1, JFEM_ERR_UNSUPPORTED, "polynomial degree must be 1..4");
    for (int a = 0; a < g.d; a++) {
        cells_1d(g.start[a], g.end[a], g.n[a], C.s1d[a], C.e1d[a]);
        std::map<std::string, int> seen;
        C.cls[a].resize(g.n[a]);
        for (int j = 0; j < g.n[a]; j++) {
            Tab1D t = make_tab(C.s1d[a][j], C.e1d[a][j], g.k, qp, qw, C.Q);
            std::string key(reinterpret_cast<const char*>(&t), sizeof(t));
            auto it = seen.find(key);
            if (it == seen.end()) {
                const int id = (int)C.tabs[a].size();
                seen.emplace(key, id);
                C.tabs[a].push_back(t);
                C.cls[a][j] = id;
            } else {
                C.cls[a][j] = it->second;
            }
        }
        C.ncls[a] = (int)C.tabs[a].size();
    }
    return C;
}

// 1D element matrices of the first cell of axis a, from the same quadrature table the assembly
// uses: K1[i][j] = sum_q w_q phi_i'(q) phi_j'(q), M1[i][j] = sum_q w_q phi_i(q) phi_j(q), (k+1) x (k+1)
// row-major.  Tensor-product Gauss quadrature factorises over the axes, so a cell's GRAD_GRAD
// matrix is sum_a K1_a (x) M1_b (x) ... and its VALUE_VALUE matrix the product of the M1_a (the
// cells of an axis differ in length by rounding only).
void element_1d(const Grid& g, int a, int rule, double* K1, double* M1)
{
    double qp[8], qw[8];
    const int Q = quad_rule(rule, qp, qw);
    JFEM_CHECK(Q > 0, JFEM_ERR_ARG, "unknown quadrature rule");
    JFEM_CHECK(g.k >= 1 && g.k < MAXK1, JFEM_ERR_UNSUPPORTED, "polynomial degree must be 1..4");
    std::vector<double> s1, e1;
    cells_1d(g.start[a], g.end[a], g.n[a], s1, e1);
    const Tab1D t = make_tab(s1[0], e1[0], g.k, qp, qw, Q);
    const int N = g.k + 1;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            double kk = 0, mm = 0;
            for (int q = 0; q < Q; q++) {
                kk += t.w[q] * t.der[i][q] * t.der[j][q];
                mm += t.w[q] * t.val[i][q] * t.val[j][q];
            }
            K1[i * N + j] = kk;
            M1[i * N + j] = mm;
        }
}

struct DevTables {
    Tab1D* tabs[MAXD] = {nullptr, nullptr, nullptr};
    int* cls[MAXD] = {nullptr, nullptr, nullptr};
    ~DevTables()
    {
        for (int a = 0; a < MAXD; a++) { (void)hipFree(tabs[a]); (void)hipFree(cls[a]); }
    }
};

static void upload_classes(jfem_ctx* c, const Grid& g, const ClassTables& C, DevTables& D)
{
    for (int a = 0; a < g.d; a++) {
        JFEM_HIP(hipMalloc(&D.tabs[a], sizeof(Tab1D) * C.tabs[a].size()));
        JFEM_HIP(hipMemcpyAsync(D.tabs[a], C.tabs[a].data(), sizeof(Tab1D) * C.tabs[a].size(), hipMemcpyHostToDevice, c->stream));
        JFEM_HIP(hipMalloc(&D.cls[a], sizeof(int) * g.n[a]));
        JFEM_HIP(hipMemcpyAsync(D.cls[a], C.cls[a].data(), sizeof(int) * g.n[a], hipMemcpyHostToDevice, c->stream));
    }
}

int choose_lanes(double mean_row)
{
    if (mean_row <= 6) return 4;
    if (mean_row <= 12) return 8;
    if (mean_row <= 24) return 16;
    if (mean_row <= 48) return 32;
    return 64;
}

// Candidate tables of k_asm_cells for dimension d and degree k: for each combination of per-axis
// row types (type = 2 [lo > 0] + [hi > lo], two bits per axis), fwd[t] = the offsets of candidate
// t (4 bits per axis) and inv[offsets] = t.  The candidate order is the one decode_candidate
// produces: per axis the segments of the owner cells lo - 1 (the point k*lo, when lo > 0), lo and
// hi, the segment of axis 0 most significant, then inside the owner box axis 0 fastest.
static void cell_asm_tables(int d, int k, std::vector<unsigned short>& fwd, std::vector<unsigned char>& inv)
{
    const int W1 = 2 * k + 1;
    int wmax = 1;
    for (int a = 0; a < d; a++) wmax *= W1;
    const int ncombo = 1 << (2 * d);
    fwd.assign((size_t)ncombo * wmax, 0);
    inv.assign((size_t)ncombo * wmax, 255);
    for (int combo = 0; combo < ncombo; combo++) {
        int nseg[MAXD], first[MAXD][3], cnt[MAXD][3], W[MAXD];
        for (int a = 0; a < d; a++) {
            const int tau = (combo >> (2 * a)) & 3;
            const bool lo_pos = tau >> 1, hi_gt = tau & 1;
            if (lo_pos) {
                const int f[3] = {0, 1, k + 1}, c[3] = {1, k, k};
                nseg[a] = hi_gt ? 3 : 2;
                for (int s = 0; s < 3; s++) { first[a][s] = f[s]; cnt[a][s] = c[s]; }
            } else {
                const int f[3] = {0, k + 1, 0}, c[3] = {k + 1, k, 0};
                nseg[a] = hi_gt ? 2 : 1;
                for (int s = 0; s < 3; s++) { first[a][s] = f[s]; cnt[a][s] = c[s]; }
            }
            W[a] = 0;
            for (int s = 0; s < nseg[a]; s++) W[a] += cnt[a][s];
        }
        int tailW[MAXD], tot = 1;
        for (int a = d - 1; a >= 0; a--) { tailW[a] = tot; tot *= W[a]; }
        for (int t = 0; t < tot; t++) {
            int rem = t, M = 1, cs[MAXD], fs[MAXD];
            for (int a = 0; a < d; a++) {
                int s = 0;
                while (s < nseg[a] - 1 && rem >= M * cnt[a][s] * tailW[a]) { rem -= M * cnt[a][s] * tailW[a]; s++; }
                cs[a] = cnt[a][s];
                fs[a] = first[a][s];
                M *= cs[a];
            }
            unsigned packed = 0;
            int oidx = 0, wm = 1;
            for (int a = 0; a < d; a++) {
                const int o = fs[a] + rem % cs[a];
                rem /= cs[a];
                packed |= (unsigned)o << (4 * a);
                oidx += o * wm;
                wm *= W1;
            }
            fwd[(size_t)combo * wmax + t] = (unsigned short)packed;
            inv[(size_t)combo * wmax + oidx] = (unsigned char)t;
        }
    }
}

// Column offsets of k_asm_cells' inner rows: delta[cls * wmax + t] = dof(q_t) - dof(p) for a
// representative row p of each class combination (per axis 0: a vertex between two cells, 1: a
// node inside a cell, 2: the last vertex; their cells lo, lo + 1 with index >= 2).  Returns
// false (no inner rows) when an axis has fewer than 4 cells.
static bool cell_asm_delta(const Grid& g, const std::vector<unsigned short>& fwd, std::vector<int>& delta)
{
    const int d = g.d, k = g.k, W1 = 2 * k + 1;
    int wmax = 1, ncls = 1;
    for (int a = 0; a < d; a++) { wmax *= W1; ncls *= 3; }
    delta.assign((size_t)ncls * wmax + 64, 0);
    for (int a = 0; a < d; a++)
        if (g.n[a] < 4) return false;
    const GridDev gd = to_dev(g);
    for (int cl = 0; cl < ncls; cl++) {
        int p[MAXD] = {0, 0, 0}, lo[MAXD] = {0, 0, 0}, combo = 0, total = 1, r = cl;
        bool ok = true;
        for (int a = 0; a < d; a++) {
            const int ca = r % 3;
            r /= 3;
            int nc;
            if (ca == 0) { p[a] = 3 * k; lo[a] = 2; nc = 2; }
            else if (ca == 1) { p[a] = 2 * k + 1; lo[a] = 2; nc = 1; ok = ok && k >= 2; }
            else { p[a] = k * g.n[a]; lo[a] = g.n[a] - 1; nc = 1; }
            combo |= (nc == 2 ? 3 : 2) << (2 * a);
            total *= k * nc + 1;
        }
        if (!ok) continue;   // no nodes inside a cell for k = 1
        const long row = dof_of_lattice(gd, p);
        for (int t = 0; t < total; t++) {
            const unsigned f = fwd[(size_t)combo * wmax + t];
            int q[MAXD] = {0, 0, 0};
            for (int a = 0; a < d; a++) q[a] = k * lo[a] + (int)((f >> (4 * a)) & 15u);
            delta[(size_t)cl * wmax + t] = (int)(dof_of_lattice(gd, q) - row);
        }
    }
    return true;
}

jfem_mat* assemble_matrix(jfem_space* s, int nint, const int* kinds, const double* weights, const double* coeff_host,
                          int rule)
{
    jfem_ctx* c = s->ctx;
    const Grid& g = s->grid;
    JFEM_CHECK(nint >= 1 && nint <= 4, JFEM_ERR_ARG, "1..4 integrals supported");
    bool full = false;
    for (int i = 0; i < nint; i++) {
        JFEM_CHECK(kinds[i] >= 0 && kinds[i] <= 4, JFEM_ERR_UNSUPPORTED, "unknown integral name");
        if (kinds[i] >= 3) full = true;
        // every kind but GRAD_GRAD evaluates value(x) / gradient(x) of the whole shape function, i.e.
        // in the first of its support cells containing x (M/basic/ScalarShapeFunction.java:26-39), not
        // in the integrated cell (TPCellIntegral.java:124-157).  The two differ only at points on the
        // cell's faces: there the neighbour's normal derivative differs, and its values differ by
        // rounding, which decides whether a lumped (exactly zero) entry enters the pattern.
        // Gauss-Lobatto points lie on the faces, so those kinds need a Gauss rule.
        JFEM_CHECK(kinds[i] == JFEM_GRAD_GRAD || (rule >= 1 && rule <= 5), JFEM_ERR_UNSUPPORTED,
                   "VALUE_VALUE / H1 / GRAD_VALUE / VALUE_GRAD need a Gauss rule (Gauss-Lobatto points lie on "
                   "cell faces, where the reference evaluates the first support cell containing the point)");
    }
    ClassTables C = build_classes(g, rule);
    DevTables D;
    upload_classes(c, g, C, D);
    ElemParams EP;
    std::memset(&EP, 0, sizeof(EP));
    EP.d = g.d; EP.k = g.k; EP.Q = C.Q; EP.nloc = g.nloc(); EP.nint = nint;
    EP.ntuples = 1;
    for (int a = 0; a < g.d; a++) { EP.ncls[a] = C.ncls[a]; EP.ntuples *= C.ncls[a]; }
    // weights are packed: one value per scalar-weighted integral, dim values per vector-weighted one
    EP.full = full ? 1 : 0;
    for (int i = 0, o = 0; i < nint; i++) {
        EP.kinds[i] = kinds[i];
        if (kinds[i] >= 3) {
            for (int a = 0; a < g.d; a++) EP.wvec[i][a] = weights[o + a];
            o += g.d;
        } else {
            EP.wconst[i] = weights[o++];
        }
    }
    const long Esize = (long)nint * EP.ntuples * EP.nloc * EP.nloc;
    JFEM_CHECK(Esize < (1l << 30), JFEM_ERR_UNSUPPORTED, "too many distinct cell geometries");
    const long n = g.ndofs();
    // node pairs iv <= iu (all pairs for non-symmetric integrals), packed iv << 16 | iu; the
    // host vectors of asynchronous uploads live until the function's last synchronisation
    std::vector<int> pairs;
    for (int v = 0; v < EP.nloc; v++)
        for (int u = full ? 0 : v; u < EP.nloc; u++) pairs.push_back(v << 16 | u);
    // every temporary in the context's set-up scratch (no allocation or free per call once it
    // has grown to this grid); 64 doubles of padding after E: k_asm_cells reads whole
    // wavefronts past a table row's end
    ScratchPlan plan;
    const size_t o_E = plan.add(sizeof(double) * (Esize + 64)), o_pairs = plan.add(sizeof(int) * pairs.size());
    const size_t o_emin = plan.add(sizeof(unsigned long long)), o_bad = plan.add(sizeof(int));
    const size_t o_coeff = plan.add(coeff_host ? sizeof(double) * g.ncells() : 0);
    const size_t o_counts = plan.add(sizeof(int) * (n > 0 ? n : 1)), o_scan = plan.add(scan_work_bytes(n));
    unsigned char* ws = static_cast<unsigned char*>(ctx_scratch(c, SCRATCH_ASM, plan.total));
    double* E = reinterpret_cast<double*>(ws + o_E);
    int* pairs_d = reinterpret_cast<int*>(ws + o_pairs);
    unsigned long long* emin_d = reinterpret_cast<unsigned long long*>(ws + o_emin);
    int* bad = reinterpret_cast<int*>(ws + o_bad);
    double* coeff = coeff_host ? reinterpret_cast<double*>(ws + o_coeff) : nullptr;
    int* counts = reinterpret_cast<int*>(ws + o_counts);
    long* scan_ws = reinterpret_cast<long*>(ws + o_scan);
    JFEM_HIP(hipMemsetAsync(E + Esize, 0, sizeof(double) * 64, c->stream));
    {
        JFEM_HIP(hipMemcpyAsync(pairs_d, pairs.data(), sizeof(int) * pairs.size(), hipMemcpyHostToDevice, c->stream));
        const long work = (long)pairs.size() * EP.ntuples * nint;
        const int eg = grid_for(work, 256, 16384);
        const int key = (g.d * 10 + g.k) * 10 + C.Q;
        switch (key) {
        case 325: k_element_tables<3, 2, 5><<<eg, 256, 0, c->stream>>>(EP, D.tabs[0], D.tabs[1], D.tabs[2], pairs_d, E); break;
        case 315: k_element_tables<3, 1, 5><<<eg, 256, 0, c->stream>>>(EP, D.tabs[0], D.tabs[1], D.tabs[2], pairs_d, E); break;
        case 215: k_element_tables<2, 1, 5><<<eg, 256, 0, c->stream>>>(EP, D.tabs[0], D.tabs[1], D.tabs[2], pairs_d, E); break;
        case 225: k_element_tables<2, 2, 5><<<eg, 256, 0, c->stream>>>(EP, D.tabs[0], D.tabs[1], D.tabs[2], pairs_d, E); break;
        default: k_element_tables<0, 0, 0><<<eg, 256, 0, c->stream>>>(EP, D.tabs[0], D.tabs[1], D.tabs[2], pairs_d, E); break;
        }
        JFEM_HIP(hipGetLastError());
    }
    // the smallest |entry| of the element tables, read back (into pinned memory) after the host
    // work below
    const bool one_integral = nint == 1;
    if (one_integral) {
        JFEM_HIP(hipMemsetAsync(emin_d, 0xff, sizeof(unsigned long long), c->stream));
        k_absmin<<<grid_for(Esize, 256, 256), 256, 0, c->stream>>>(Esize, E, emin_d);
        JFEM_HIP(hipGetLastError());
    }
    if (coeff_host)
        JFEM_HIP(hipMemcpyAsync(coeff, coeff_host, sizeof(double) * g.ncells(), hipMemcpyHostToDevice, c->stream));
    RowParams RP;
    std::memset(&RP, 0, sizeof(RP));
    RP.g = to_dev(g);
    RP.nloc = EP.nloc; RP.nint = nint; RP.ntuples = EP.ntuples; RP.E = E; RP.coeff = coeff;
    for (int a = 0; a < MAXD; a++) { RP.cls[a] = D.cls[a]; RP.ncls[a] = a < g.d ? C.ncls[a] : 1; }

    auto* A = new jfem_mat();
    A->ctx = c;
    A->rows = A->cols = n;
    JFEM_HIP(hipMalloc(&A->rowptr, sizeof(int) * (n + 1)));
    int maxW = 1;
    for (int a = 0; a < g.d; a++) maxW *= 2 * g.k + 1;
    const int G = maxW <= 16 ? 16 : (maxW <= 32 ? 32 : 64);
    const int groups_per_block = 256 / G;
    const int blocks = grid_for(n, groups_per_block, 65536);
    JFEM_HIP(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    // tables in LDS when they fit in 48 KiB (any uniform or few-class grid)
    long cls_ints = 0;
    for (int a = 0; a < g.d; a++) cls_ints += g.n[a];
    const size_t lds_bytes = sizeof(double) * Esize + sizeof(int) * cls_ints;
    const bool use_lds = lds_bytes <= 48 * 1024;
    auto run = [&](auto kern) {
        kern<<<blocks, 256, use_lds ? lds_bytes : 0, c->stream>>>(RP, n, counts, A->rowptr, A->colidx, A->vals, bad);
        JFEM_HIP(hipGetLastError());
    };
    auto run2 = [&](auto kern_lds, auto kern_glob) { use_lds ? run(kern_lds) : run(kern_glob); };
    // k <= 2: the cell-major kernel, on a grid of co-resident blocks (each block stages the
    // candidate tables once); element tables and cell classes in LDS when they fit in 16 KiB
    const bool cellwise = g.k <= 2;
    std::vector<unsigned short> fwd_h;
    std::vector<unsigned char> inv_h;
    std::vector<int> delta_h;
    unsigned short* fwd_d = nullptr;
    unsigned char* inv_d = nullptr;
    int* delta_d = nullptr;
    std::vector<AxEntry> axtab_h;
    std::vector<int> order_h;
    int* order_d = nullptr;
    AxOffsets aoff;
    AxEntry* axtab_d = nullptr;
    int inner_ok = 0;
    // the cell-major kernel's tables (host work and uploads), made once cols and vals exist:
    // the element tables and the pattern run meanwhile
    auto prepare_cells = [&]() {
        if (cellwise) {
            cell_asm_axes(g, C.cls, C.ncls, axtab_h, aoff);
            // lines (lattice points of the fixed axes, last fixed axis fastest) ordered by the
            // classes of the cells containing them, axis by axis: each fixed axis' points sorted by
            // their cells' classes, the lines their product in that order
            {
                std::vector<int> sorted[MAXD];
                long nl = 1;
                for (int a = 0; a + 1 < g.d; a++) {
                    const int Na = g.k * g.n[a] + 1;
                    nl *= Na;
                    std::vector<std::pair<long, int>> keyed((size_t)Na);
                    for (int pa = 0; pa < Na; pa++) {
                        const AxEntry& e = axtab_h[(size_t)aoff.o[a] + pa];
                        const long c0 = C.ncls[a] == 1 ? 0 : C.cls[a][e.lo];
                        const long c1 = (e.info & 3) == 2 && C.ncls[a] > 1 ? C.cls[a][e.lo + 1] : c0;
                        keyed[(size_t)pa] = {c0 * (C.ncls[a] + 1) + c1, pa};
                    }
                    std::stable_sort(keyed.begin(), keyed.end(),
                                     [](const std::pair<long, int>& x, const std::pair<long, int>& y) { return x.first < y.first; });
                    for (auto& kv : keyed) sorted[a].push_back(kv.second);
                }
                order_h.assign((size_t)nl, 0);
                if (g.d == 3) {
                    const int N1 = g.k * g.n[1] + 1;
                    size_t i = 0;
                    for (int p0 : sorted[0])
                        for (int p1 : sorted[1]) order_h[i++] = p0 * N1 + p1;
                } else if (g.d == 2) {
                    order_h = sorted[0];
                }
            }
            cell_asm_tables(g.d, g.k, fwd_h, inv_h);
            fwd_h.resize(fwd_h.size() + 64, 0);   // whole wavefronts read past the last combination's entries
            inner_ok = cell_asm_delta(g, fwd_h, delta_h) ? 1 : 0;
            ScratchPlan cp;
            const size_t o_order = cp.add(sizeof(int) * order_h.size()), o_ax = cp.add(sizeof(AxEntry) * axtab_h.size());
            const size_t o_fwd = cp.add(sizeof(unsigned short) * fwd_h.size()), o_inv = cp.add(inv_h.size());
            const size_t o_delta = cp.add(sizeof(int) * (delta_h.size() > 0 ? delta_h.size() : 1));
            unsigned char* cs = static_cast<unsigned char*>(ctx_scratch(c, SCRATCH_CELLS, cp.total));
            order_d = reinterpret_cast<int*>(cs + o_order);
            axtab_d = reinterpret_cast<AxEntry*>(cs + o_ax);
            fwd_d = reinterpret_cast<unsigned short*>(cs + o_fwd);
            inv_d = cs + o_inv;
            delta_d = reinterpret_cast<int*>(cs + o_delta);
            JFEM_HIP(hipMemcpyAsync(order_d, order_h.data(), sizeof(int) * order_h.size(), hipMemcpyHostToDevice,
                                    c->stream));
            JFEM_HIP(hipMemcpyAsync(axtab_d, axtab_h.data(), sizeof(AxEntry) * axtab_h.size(), hipMemcpyHostToDevice,
                                    c->stream));
            JFEM_HIP(hipMemcpyAsync(fwd_d, fwd_h.data(), sizeof(unsigned short) * fwd_h.size(), hipMemcpyHostToDevice,
                                    c->stream));
            JFEM_HIP(hipMemcpyAsync(inv_d, inv_h.data(), inv_h.size(), hipMemcpyHostToDevice, c->stream));
            JFEM_HIP(hipMemcpyAsync(delta_d, delta_h.data(), sizeof(int) * delta_h.size(), hipMemcpyHostToDevice,
                                    c->stream));
        }
    };
    const size_t cell_lds_bytes = sizeof(double) * (Esize + 64);   // E with the same padding
    const bool cell_lds = cell_lds_bytes <= 16 * 1024;
    auto run_cells = [&](auto kern, int GG) {
        const size_t dyn = cell_lds ? cell_lds_bytes : 0;
        long nlines = 1;
        for (int a = 0; a + 1 < g.d; a++) nlines *= (long)g.k * g.n[a] + 1;
        // lines split into segments (a multiple of the rows per wavefront) until there are
        // about 32 wavefronts per CU
        const int rpw = 64 / GG;
        const int nl = g.k * g.n[g.d - 1] + 1;
        const long want = (long)c->cu_count * 32;
        int nseg = (int)std::min<long>((want + nlines - 1) / nlines, (nl + rpw - 1) / rpw);
        nseg = std::max(nseg, 1);
        int seg_len = ((nl + nseg - 1) / nseg + rpw - 1) / rpw * rpw;
        nseg = (nl + seg_len - 1) / seg_len;
        const long nunits = nlines * nseg;
        const long nb = (nunits + 3) / 4;
        JFEM_CHECK(nb < (1l << 31), JFEM_ERR_UNSUPPORTED, "too many lattice lines");
        kern<<<(int)nb, 256, dyn, c->stream>>>(RP, nunits, nseg, seg_len, counts, A->rowptr, A->colidx, A->vals,
                                               bad, fwd_d, inv_d, delta_d, axtab_d, aoff, order_d, inner_ok);
        JFEM_HIP(hipGetLastError());
    };
    auto run_cells4 = [&](auto kern_lds_c, auto kern_lds, auto kern_glob_c, auto kern_glob, int GG) {
        if (cell_lds) coeff ? run_cells(kern_lds_c, GG) : run_cells(kern_lds, GG);
        else coeff ? run_cells(kern_glob_c, GG) : run_cells(kern_glob, GG);
    };
#define JFEM_ASM_CASE(DD, KK, GG)                                                                   \
    fill ? run2(k_assemble_rows<DD, KK, GG, true, true>, k_assemble_rows<DD, KK, GG, true, false>)  \
         : run2(k_assemble_rows<DD, KK, GG, false, true>, k_assemble_rows<DD, KK, GG, false, false>)
#define JFEM_CELL_CASE(DD, KK, GG)                                                                           \
    fill ? run_cells4(k_asm_cells<DD, KK, GG, true, true, true>, k_asm_cells<DD, KK, GG, true, true, false>,     \
                      k_asm_cells<DD, KK, GG, true, false, true>, k_asm_cells<DD, KK, GG, true, false, false>, GG) \
         : run_cells4(k_asm_cells<DD, KK, GG, false, true, true>, k_asm_cells<DD, KK, GG, false, true, false>,   \
                      k_asm_cells<DD, KK, GG, false, false, true>, k_asm_cells<DD, KK, GG, false, false, false>, GG)
    // k_assemble_rows: G follows from (d, k), (2k+1)^d candidates per row; k_asm_cells: the
    // smallest power of two >= (k+1)^d lanes per row
    auto launch = [&](bool fill) {
        const int dk = g.d * 10 + g.k;
        switch (dk) {
        case 11: JFEM_CELL_CASE(1, 1, 4); break;
        case 12: JFEM_CELL_CASE(1, 2, 4); break;
        case 13: JFEM_ASM_CASE(1, 3, 16); break;
        case 21: JFEM_CELL_CASE(2, 1, 4); break;
        case 22: JFEM_CELL_CASE(2, 2, 16); break;
        case 23: JFEM_ASM_CASE(2, 3, 64); break;
        case 31: JFEM_CELL_CASE(3, 1, 8); break;
        case 32: JFEM_CELL_CASE(3, 2, 64); break;
        case 33: JFEM_ASM_CASE(3, 3, 64); break;
        default:
            if (G == 16) JFEM_ASM_CASE(0, 0, 16);
            else if (G == 32) JFEM_ASM_CASE(0, 0, 32);
            else JFEM_ASM_CASE(0, 0, 64);
        }
    };
#undef JFEM_ASM_CASE
#undef JFEM_CELL_CASE
    // With one integral, no exactly-zero element entry, and coefficients whose products with the
    // entries cannot round to zero, every cell contribution is nonzero, so every structural
    // candidate is in the pattern and the row lengths have a closed form; otherwise the count
    // pass evaluates them.  The element tables' smallest entry is known only on the device, so
    // with one integral and nonzero coefficients the structural pattern is assumed: its row
    // lengths, row pointer, SpMV blocks and the fill are queued at once, and the whole assembly
    // waits once.  Should the tables hold an exact zero, the count pass and the fill run again.
    // The entries of the structural pattern bound nnz (the pattern is a subset of it): cols and
    // vals are allocated for it while the device works.
    long nnz_bound = 1;
    for (int a = 0; a < g.d; a++) {
        long sa = 0;
        for (int pa = 0; pa <= g.k * g.n[a]; pa++) {
            int l, h;
            if (pa % g.k == 0) { l = pa / g.k - 1; h = pa / g.k; } else { l = h = pa / g.k; }
            l = std::max(l, 0);
            h = std::min(h, g.n[a] - 1);
            sa += (long)g.k * (h - l + 1) + 1;
        }
        nnz_bound *= sa;
    }
    JFEM_CHECK(nnz_bound < (1l << 31) - 1, JFEM_ERR_UNSUPPORTED, "matrix has 2^31 or more entries (int32 CSR indices)");
    double cmin = 1.0;
    if (coeff_host)
        for (long i = 0; i < g.ncells(); i++) cmin = std::min(cmin, std::fabs(coeff_host[i]));
    const long tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    long* tail = c->pin + PIN_LONGS - 2;   // the fill's status and the scan's total
    // row lengths (closed form or count pass), row pointer, SpMV blocks; the fill is launched by
    // the caller once cols and vals exist
    // (the closed form reads neither the element tables nor anything else queued on the
    // context's stream, so it runs on the auxiliary stream beside the element tables; the
    // context's stream waits for it before the fill)
    auto pattern = [&](bool structural) {
        hipStream_t ps = structural ? c->aux : c->stream;
        if (structural) {
            const int cb = grid_for(n, 256, 65536);
            switch (g.d * 10 + g.k) {
            case 11: k_count_structural<1, 1><<<cb, 256, 0, ps>>>(RP.g, n, counts); break;
            case 12: k_count_structural<1, 2><<<cb, 256, 0, ps>>>(RP.g, n, counts); break;
            case 21: k_count_structural<2, 1><<<cb, 256, 0, ps>>>(RP.g, n, counts); break;
            case 22: k_count_structural<2, 2><<<cb, 256, 0, ps>>>(RP.g, n, counts); break;
            case 31: k_count_structural<3, 1><<<cb, 256, 0, ps>>>(RP.g, n, counts); break;
            case 32: k_count_structural<3, 2><<<cb, 256, 0, ps>>>(RP.g, n, counts); break;
            default: k_count_structural<0, 0><<<cb, 256, 0, ps>>>(RP.g, n, counts); break;
            }
            JFEM_HIP(hipGetLastError());
        } else {
            launch(false);
        }
        // with the structural pattern the entry count is nnz_bound and nothing is read back; the
        // scan's total comes back with the fill's status and is checked then
        A->nnz = scan_counts(c, counts, n, A->rowptr, scan_ws, structural ? nnz_bound : -1, ps);
        JFEM_CHECK(A->nnz <= nnz_bound, JFEM_ERR_STATE, "more entries than the structural pattern");
        mat_blocks_queue(A, ps);
        if (structural) {
            JFEM_HIP(hipEventRecord(c->aux_done, c->aux));
            JFEM_HIP(hipStreamWaitEvent(c->stream, c->aux_done, 0));
        }
    };
    auto fill_and_status = [&]() {
        launch(true);
        JFEM_HIP(hipMemcpyAsync(tail, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        JFEM_HIP(hipMemcpyAsync(tail + 1, scan_ws + (n > 0 ? n : 1) + tiles + 1, sizeof(long), hipMemcpyDeviceToHost,
                                c->stream));
    };
    const bool assume = one_integral && cmin > 0;
    if (assume) pattern(true);
    JFEM_HIP(hipMalloc(&A->colidx, sizeof(int) * (nnz_bound + NNZ_PAD)));
    JFEM_HIP(hipMalloc(&A->vals, sizeof(double) * (nnz_bound + NNZ_PAD)));
    prepare_cells();
    if (assume) fill_and_status();
    if (one_integral) JFEM_HIP(hipMemcpyAsync(c->pin, emin_d, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    JFEM_HIP(hipStreamSynchronize(c->stream));
    bool structural = one_integral;
    if (structural) {
        double emin;
        std::memcpy(&emin, c->pin, sizeof(emin));
        structural = emin > 0 && cmin > 0 && emin * cmin > 1e-300;
    }
    if (!structural) {
        // the row lengths need the count pass (and the fill its row pointer)
        JFEM_HIP(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
        pattern(false);
        fill_and_status();
        JFEM_HIP(hipStreamSynchronize(c->stream));
    }
    const int bad_h = *reinterpret_cast<const int*>(tail);
    const long total_h = tail[1];
    if (bad_h || total_h != A->nnz) {
        mat_free(A);
        JFEM_CHECK(false, JFEM_ERR_STATE, "assembled rows differ from their reserved lengths");
    }
    A->lanes = choose_lanes(n > 0 ? (double)A->nnz / n : 1.0);
    mat_blocks_done(A);
    return A;
}

void assemble_rhs_sines(jfem_space* s, int nterms, const int* kw, const double* amp, int rule, double* rhs,
                        const int* cosine)
{
    jfem_ctx* c = s->ctx;
    const Grid& g = s->grid;
    JFEM_CHECK(nterms >= 1, JFEM_ERR_ARG, "no source terms");
    // TPRightHandSideIntegral VALUE evaluates shapeFunction.value(x) (TPRightHandSideIntegral.java:26-29),
    // the first support cell containing x: per-cell tables model it at interior (Gauss) points only
    JFEM_CHECK(rule >= 1 && rule <= 5, JFEM_ERR_UNSUPPORTED,
               "right-hand side integrals need a Gauss rule (Gauss-Lobatto points lie on cell faces)");
    ClassTables C = build_classes(g, rule);
    DevTables D;
    upload_classes(c, g, C, D);
    double qp[8], qw[8];
    quad_rule(rule, qp, qw);
    // per-axis tables sin(kw * pi * x_q) (cos where the cosine flag is set); the amplitude is folded into axis 0
    DevBuf<double> dsin[MAXD];
    RhsParams P;
    std::memset(&P, 0, sizeof(P));
    P.g = to_dev(g);
    P.Q = C.Q; P.nloc = g.nloc(); P.nterms = nterms;
    for (int a = 0; a < g.d; a++) {
        std::vector<double> t((size_t)nterms * g.n[a] * C.Q);
        for (int m = 0; m < nterms; m++)
            for (int j = 0; j < g.n[a]; j++)
                for (int q = 0; q < C.Q; q++) {
#pragma clang fp contract(off)
                    const double len = C.e1d[a][j] - C.s1d[a][j];
                    const double x = qp[q] * len + C.s1d[a][j];
                    double v = cosine && cosine[m * g.d + a] ? std::cos(kw[m * g.d + a] * M_PI * x)
                                                             : std::sin(kw[m * g.d + a] * M_PI * x);
                    if (a == 0) v = amp[m] * v;
                    t[((size_t)m * g.n[a] + j) * C.Q + q] = v;
                }
        dsin[a].alloc(t.size());
        JFEM_HIP(hipMemcpyAsync(dsin[a], t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, c->stream));
        JFEM_HIP(hipStreamSynchronize(c->stream));   // t leaves scope
        P.sinq[a] = dsin[a];
        P.cls[a] = D.cls[a];
        P.ncls[a] = C.ncls[a];
    }
    vec_fill(c, rhs, 0.0, g.ndofs());
    for (int color = 0; color < (1 << g.d); color++) {
        long ncc = 1;
        bool empty = false;
        for (int a = 0; a < g.d; a++) {
            const long h = (g.n[a] - ((color >> a) & 1) + 1) / 2;
            if (h <= 0) empty = true;
            ncc *= h;
        }
        if (empty) continue;
        const int cpb = RHS_BS / P.nloc;
        int nq = 1;
        for (int a = 0; a < g.d; a++) nq *= C.Q;
        const size_t lds = sizeof(double) * cpb * nq;
        const int blocks = grid_for((ncc + cpb - 1) / cpb, 1, 65536);
        const int key = (g.d * 10 + g.k) * 10 + C.Q;
        switch (key) {
        case 325: k_rhs_color<3, 2, 5><<<blocks, RHS_BS, lds, c->stream>>>(P, D.tabs[0], D.tabs[1], D.tabs[2], color, ncc, rhs); break;
        case 315: k_rhs_color<3, 1, 5><<<blocks, RHS_BS, lds, c->stream>>>(P, D.tabs[0], D.tabs[1], D.tabs[2], color, ncc, rhs); break;
        case 215: k_rhs_color<2, 1, 5><<<blocks, RHS_BS, lds, c->stream>>>(P, D.tabs[0], D.tabs[1], D.tabs[2], color, ncc, rhs); break;
        case 225: k_rhs_color<2, 2, 5><<<blocks, RHS_BS, lds, c->stream>>>(P, D.tabs[0], D.tabs[1], D.tabs[2], color, ncc, rhs); break;
        default: k_rhs_color<0, 0, 0><<<blocks, RHS_BS, lds, c->stream>>>(P, D.tabs[0], D.tabs[1], D.tabs[2], color, ncc, rhs); break;
        }
        JFEM_HIP(hipGetLastError());
    }
    JFEM_HIP(hipStreamSynchronize(c->stream));
}

void space_boundary_mask(jfem_space* s, unsigned char* mask_dev)
{
    const long n = s->grid.ndofs();
    k_boundary_mask<<<grid_for(n, 256, 16384), 256, 0, s->ctx->stream>>>(to_dev(s->grid), n, mask_dev);
    JFEM_HIP(hipGetLastError());
}


// ------------------------- Taylor-Hood mixed system -------------------------
// TaylorHoodSpace (M/mixed/TaylorHoodSpace.java:17-92): pressure Qk and
// velocity Q(k+1)^d on one Cartesian grid.  The shape functions live in a
// TreeSet ordered by QkQkFunction.compareTo (M/mixed/QkQkFunction.java:82-93):
// velocity before pressure; velocity component-major
// (SingleComponentVectorShapeFunction.compareTo, M/basic/SingleComponentVectorShapeFunction.java:168-173);
// inside a component, and for the pressure, by node point under
// CoordinateComparator (first coordinate most significant, tolerance 1e-10,
// M/linalg/CoordinateComparator.java:15-27; ContinuousTPShapeFunction.compareTo
// :232-242).  Node spacings far above 1e-10 make that the lexicographic
// lattice order with the last axis fastest: lex(p) = sum_a p_a prod_{b>a} N_b.

__host__ __device__ inline long lex_of_lattice(int d, const int* N, const int* p)
{
    long r = 0;
    for (int a = 0; a < d; a++) r = r * N[a] + p[a];
    return r;
}

__host__ __device__ inline void lattice_of_lex(int d, const int* N, long r, int* p)
{
    for (int a = d - 1; a >= 0; a--) { p[a] = (int)(r % N[a]); r /= N[a]; }
}

// lexicographic rank of every DOF of a scalar space (DOF order -> node order)
__global__ void k_lex_rank(GridDev g, long n, int* __restrict__ out)
{
    int N[MAXD];
    for (int a = 0; a < MAXD; a++) N[a] = a < g.d ? g.k * g.n[a] + 1 : 1;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int p[MAXD] = {0, 0, 0};
        lattice_of_dof(g, i, p);
        out[i] = (int)lex_of_lattice(g.d, N, p);
    }
}

void space_lex_rank(jfem_space* s, int* out_dev)
{
    const long n = s->grid.ndofs();
    JFEM_CHECK(n < (1l << 31), JFEM_ERR_UNSUPPORTED, "2^31 or more nodes");
    if (n) k_lex_rank<<<grid_for(n, 256, 16384), 256, 0, s->ctx->stream>>>(to_dev(s->grid), n, out_dev);
    JFEM_HIP(hipGetLastError());
}

struct ThParams {
    int kv, kp, nlocv, nlocp;
    int n[MAXD];
    int NV[MAXD], NP[MAXD];   // velocity and pressure nodes per axis
    long nv, np;              // nodes of one velocity component, pressure nodes
    int ncls[MAXD];
    const int* cls[MAXD];     // class of every 1D cell per axis
    const double* Evv;        // [tuple][iv][iu], nlocv^2 per tuple (null: no velocity integral)
    const double* EB;         // [tuple][c][iv][ip], d*nlocv*nlocp per tuple (null: no DIV_VALUE)
};

// cells of one axis that contain lattice coordinate x of a degree-k lattice
__device__ inline void cells_of(int x, int k, int n, int& lo, int& hi)
{
    if (x % k == 0) { lo = x / k - 1; hi = x / k; } else { lo = hi = x / k; }
    if (lo < 0) lo = 0;
    if (hi > n - 1) hi = n - 1;
}

// entry between velocity node p (component c) and a node q of the lattice of
// degree kq: sum over the cells containing both, lexicographic cell order,
// of the element values that are not exactly zero (MatrixFESpace.addToMatrix
// skips those).  vv: q is a velocity node of the same component; otherwise q
// is a pressure node.
template <int DD>
__device__ inline double th_entry(const ThParams& P, int c, const int* p, const int* q, bool vv, bool& present)
{
#pragma clang fp contract(off)
    int lo[MAXD] = {0, 0, 0}, cnt[MAXD] = {1, 1, 1};
    const int kq = vv ? P.kv : P.kp;
#pragma unroll
    for (int a = 0; a < DD; a++) {
        int plo, phi, qlo, qhi;
        cells_of(p[a], P.kv, P.n[a], plo, phi);
        cells_of(q[a], kq, P.n[a], qlo, qhi);
        lo[a] = max(plo, qlo);
        cnt[a] = min(phi, qhi) - lo[a] + 1;
    }
    double sum = 0;
    present = false;
#pragma unroll
    for (int i0 = 0; i0 < 2; i0++)
#pragma unroll
        for (int i1 = 0; i1 < 2; i1++)
#pragma unroll
            for (int i2 = 0; i2 < (DD > 2 ? 2 : 1); i2++) {
                const int ii[MAXD] = {i0, i1, i2};
                bool in = true;
#pragma unroll
                for (int a = 0; a < DD; a++) in = in && ii[a] < cnt[a];
                if (!in) continue;
                long tup = 0;
                int iv = 0, iq = 0, mv = 1, mq = 1;
#pragma unroll
                for (int a = 0; a < DD; a++) {
                    const int C = lo[a] + ii[a];
                    tup = tup * P.ncls[a] + (P.ncls[a] == 1 ? 0 : P.cls[a][C]);
                    iv += (p[a] - P.kv * C) * mv;
                    iq += (q[a] - kq * C) * mq;
                    mv *= P.kv + 1;
                    mq *= kq + 1;
                }
                const double e = vv ? P.Evv[(tup * P.nlocv + iv) * P.nlocv + iq]
                                    : P.EB[((tup * DD + c) * P.nlocv + iv) * P.nlocp + iq];
                if (e != 0) {
                    sum = e + sum;
                    present = true;
                }
            }
    return sum;
}

// one thread per row of the mixed matrix; FILL = false counts the entries,
// FILL = true writes them (and flags rows whose count differs)
template <int DD, bool FILL>
__global__ __launch_bounds__(256) void k_th_rows(ThParams P, long nrows, int* __restrict__ counts,
                                                 const int* __restrict__ rowptr, int* __restrict__ cols,
                                                 double* __restrict__ vals, int* __restrict__ bad)
{
#pragma clang fp contract(off)
    const long nvel = DD * P.nv;
    for (long row = blockIdx.x * (long)blockDim.x + threadIdx.x; row < nrows; row += (long)gridDim.x * blockDim.x) {
        int written = 0;
        const int base = FILL ? rowptr[row] : 0;
        auto emit = [&](long col, double v) {
            if (FILL) { cols[base + written] = (int)col; vals[base + written] = v; }
            written++;
        };
        int lo[MAXD] = {0, 0, 0}, hi[MAXD] = {0, 0, 0};
        if (row < nvel) {
            const int c = (int)(row / P.nv);
            int p[MAXD] = {0, 0, 0};
            lattice_of_lex(DD, P.NV, row - c * P.nv, p);
#pragma unroll
            for (int a = 0; a < DD; a++) cells_of(p[a], P.kv, P.n[a], lo[a], hi[a]);
            if (P.Evv) {
                int q[MAXD] = {0, 0, 0};
                int qlo[MAXD], qhi[MAXD];
#pragma unroll
                for (int a = 0; a < DD; a++) { qlo[a] = lo[a] * P.kv; qhi[a] = (hi[a] + 1) * P.kv; q[a] = qlo[a]; }
                for (;;) {
                    bool present;
                    const double v = th_entry<DD>(P, c, p, q, true, present);
                    const long col = c * P.nv + lex_of_lattice(DD, P.NV, q);
                    if (row == 0 && col == 0) present = true;   // dummy (0,0) entry of a new SparseMatrix
                    if (present) emit(col, v);
                    int a = DD - 1;
                    while (a >= 0 && q[a] == qhi[a]) { q[a] = qlo[a]; a--; }
                    if (a < 0) break;
                    q[a]++;
                }
            } else if (row == 0) {
                emit(0, 0.0);
            }
            if (P.EB) {
                int q[MAXD] = {0, 0, 0};
                int qlo[MAXD], qhi[MAXD];
#pragma unroll
                for (int a = 0; a < DD; a++) { qlo[a] = lo[a] * P.kp; qhi[a] = (hi[a] + 1) * P.kp; q[a] = qlo[a]; }
                for (;;) {
                    bool present;
                    const double v = th_entry<DD>(P, c, p, q, false, present);
                    if (present) emit(nvel + lex_of_lattice(DD, P.NP, q), v);
                    int a = DD - 1;
                    while (a >= 0 && q[a] == qhi[a]) { q[a] = qlo[a]; a--; }
                    if (a < 0) break;
                    q[a]++;
                }
            }
        } else if (P.EB) {
            int q[MAXD] = {0, 0, 0};
            lattice_of_lex(DD, P.NP, row - nvel, q);
#pragma unroll
            for (int a = 0; a < DD; a++) cells_of(q[a], P.kp, P.n[a], lo[a], hi[a]);
            for (int c = 0; c < DD; c++) {
                int p[MAXD] = {0, 0, 0};
                int plo[MAXD], phi[MAXD];
#pragma unroll
                for (int a = 0; a < DD; a++) { plo[a] = lo[a] * P.kv; phi[a] = (hi[a] + 1) * P.kv; p[a] = plo[a]; }
                for (;;) {
                    bool present;
                    const double v = th_entry<DD>(P, c, p, q, false, present);
                    if (present) emit(c * P.nv + lex_of_lattice(DD, P.NV, p), v);
                    int a = DD - 1;
                    while (a >= 0 && p[a] == phi[a]) { p[a] = plo[a]; a--; }
                    if (a < 0) break;
                    p[a]++;
                }
            }
        }
        if (!FILL) counts[row] = written;
        else if (written != rowptr[row + 1] - base) atomicOr(bad, 1);
    }
}

// DIV_VALUE element values (MixedTPCellIntegral.evaluatePressureVelocityIntegral,
// M/mixed/MixedTPCellIntegral.java:37-59): for velocity function (component c,
// local node iv) and pressure function ip, integrateNonTensorProduct
// (M/tensorproduct/TPCellIntegral.java:78-102) of
//   divergence * value * weight,  divergence = gradient component c of the
// scalar component function (SingleComponentVectorShapeFunction.divergence
// :116-120, fastGradientInCell product order, ContinuousTPShapeFunction.java:173-198),
// value = product of the 1D values (fastValueInCell :154-171), then times each
// axis' quadrature weight, summed over the points with the last axis fastest.
// VALUE_GRAD (:50-57) is velocity value . pressure gradient * weight; the velocity value vector
// has the component function's value at entry c and exact zeros elsewhere, so the inner product
// (DenseVector.inner, M/linalg/DenseVector.java:104-112) is value * gradient component c exactly.
// With both integrals the element value is their sum, as writeCellIntegralsToMatrix adds them.
#pragma clang fp contract(off)
static void div_value_table(int d, int kv, int kp, int Q, const Tab1D* const* tv, const Tab1D* const* tp, double w,
                            double* out, bool value_grad = false)
{
    const int nlv = (int)std::lround(std::pow(kv + 1, d)), nlp = (int)std::lround(std::pow(kp + 1, d));
    int nq = 1;
    for (int a = 0; a < d; a++) nq *= Q;
    for (int c = 0; c < d; c++)
        for (int iv = 0; iv < nlv; iv++)
            for (int ip = 0; ip < nlp; ip++) {
                int lv[MAXD] = {0, 0, 0}, lp[MAXD] = {0, 0, 0};
                int rv = iv, rp = ip;
                for (int a = 0; a < d; a++) { lv[a] = rv % (kv + 1); rv /= kv + 1; lp[a] = rp % (kp + 1); rp /= kp + 1; }
                double ret = 0;
                for (int q = 0; q < nq; q++) {
                    int qa[MAXD] = {0, 0, 0};
                    int qq = q;
                    for (int a = d - 1; a >= 0; a--) { qa[a] = qq % Q; qq /= Q; }
                    double val;
                    if (!value_grad) {
                        double div = 1;
                        for (int j = 0; j < d; j++)
                            div *= j == c ? tv[j]->der[lv[j]][qa[j]] : tv[j]->val[lv[j]][qa[j]];
                        double pv = 1;
                        for (int a = 0; a < d; a++) pv *= tp[a]->val[lp[a]][qa[a]];
                        val = div * pv * w;
                    } else {
                        double vv = 1;
                        for (int a = 0; a < d; a++) vv *= tv[a]->val[lv[a]][qa[a]];
                        double g = 1;
                        for (int j = 0; j < d; j++)
                            g *= j == c ? tp[j]->der[lp[j]][qa[j]] : tp[j]->val[lp[j]][qa[j]];
                        val = vv * g * w;
                    }
                    for (int a = 0; a < d; a++) val *= tv[a]->w[qa[a]];
                    ret += val;
                }
                out[((long)c * nlv + iv) * nlp + ip] = ret;
            }
}
#pragma clang fp contract(on)

jfem_mat* th_assemble(jfem_space* V, jfem_space* Pp, bool has_vv, double vv_w, bool has_div, double div_w, int rule,
                      bool has_vg, double vg_w)
{
    jfem_ctx* c = V->ctx;
    const Grid& gv = V->grid;
    const Grid& gp = Pp->grid;
    const int d = gv.d;
    JFEM_CHECK(d == 2 || d == 3, JFEM_ERR_UNSUPPORTED, "Taylor-Hood spaces in 2 or 3 dimensions");
    JFEM_CHECK(gp.d == d && gv.k == gp.k + 1, JFEM_ERR_ARG,
               "velocity space must have degree pressure degree + 1 on the same grid");
    for (int a = 0; a < d; a++)
        JFEM_CHECK(gv.n[a] == gp.n[a] && gv.start[a] == gp.start[a] && gv.end[a] == gp.end[a], JFEM_ERR_ARG,
                   "velocity and pressure spaces on different grids");
    JFEM_CHECK(has_vv || has_div || has_vg, JFEM_ERR_ARG, "no cell integrals");
    double qp[8], qw[8];
    const int Q = quad_rule(rule, qp, qw);
    JFEM_CHECK(Q > 0, JFEM_ERR_ARG, "unknown quadrature rule");
    // the mixed integrals evaluate value(x) / gradient(x) of the whole shape function, i.e. in the
    // first support cell containing x (ScalarShapeFunction.java:26-39); at Gauss-Lobatto points on
    // a cell face that can be the neighbouring cell, which the per-cell tables do not model
    JFEM_CHECK(rule >= 1 && rule <= 5, JFEM_ERR_UNSUPPORTED,
               "mixed integrals need a Gauss rule (Gauss-Lobatto points lie on cell faces)");
    JFEM_CHECK(gv.k < MAXK1, JFEM_ERR_UNSUPPORTED, "velocity degree must be 2..4");
    // cell classes: 1D cells with bitwise equal velocity and pressure tables
    ClassTables CV;
    CV.Q = Q;
    std::vector<Tab1D> tabp[MAXD];
    for (int a = 0; a < d; a++) {
        cells_1d(gv.start[a], gv.end[a], gv.n[a], CV.s1d[a], CV.e1d[a]);
        std::map<std::string, int> seen;
        CV.cls[a].resize(gv.n[a]);
        for (int j = 0; j < gv.n[a]; j++) {
            const Tab1D tv = make_tab(CV.s1d[a][j], CV.e1d[a][j], gv.k, qp, qw, Q);
            const Tab1D tq = make_tab(CV.s1d[a][j], CV.e1d[a][j], gp.k, qp, qw, Q);
            std::string key(reinterpret_cast<const char*>(&tv), sizeof(tv));
            key.append(reinterpret_cast<const char*>(&tq), sizeof(tq));
            auto it = seen.find(key);
            if (it == seen.end()) {
                const int id = (int)CV.tabs[a].size();
                seen.emplace(key, id);
                CV.tabs[a].push_back(tv);
                tabp[a].push_back(tq);
                CV.cls[a][j] = id;
            } else {
                CV.cls[a][j] = it->second;
            }
        }
        CV.ncls[a] = (int)CV.tabs[a].size();
    }
    long ntuples = 1;
    for (int a = 0; a < d; a++) ntuples *= CV.ncls[a];
    const int nlv = gv.nloc(), nlp = gp.nloc();
    JFEM_CHECK(ntuples * (long)nlv * nlv < (1l << 30), JFEM_ERR_UNSUPPORTED, "too many distinct cell geometries");
    DevTables D;
    upload_classes(c, gv, CV, D);
    DevBuf<double> Evv, EB;
    if (has_vv) {
        ElemParams EP;
        std::memset(&EP, 0, sizeof(EP));
        EP.d = d; EP.k = gv.k; EP.Q = Q; EP.nloc = nlv; EP.nint = 1; EP.ntuples = ntuples;
        EP.kinds[0] = 0;   // GRAD_GRAD: the Frobenius product of single-component gradients is the scalar one
        EP.wconst[0] = vv_w;
        for (int a = 0; a < d; a++) EP.ncls[a] = CV.ncls[a];
        Evv.alloc((size_t)ntuples * nlv * nlv);
        std::vector<int> pairs;
        for (int v = 0; v < nlv; v++)
            for (int u = v; u < nlv; u++) pairs.push_back(v << 16 | u);
        DevBuf<int> pairs_d(pairs.size());
        JFEM_HIP(hipMemcpyAsync(pairs_d, pairs.data(), sizeof(int) * pairs.size(), hipMemcpyHostToDevice, c->stream));
        const int eg = grid_for((long)pairs.size() * ntuples, 256, 16384);
        switch ((d * 10 + gv.k) * 10 + Q) {
        case 325: k_element_tables<3, 2, 5><<<eg, 256, 0, c->stream>>>(EP, D.tabs[0], D.tabs[1], D.tabs[2], pairs_d, Evv); break;
        case 225: k_element_tables<2, 2, 5><<<eg, 256, 0, c->stream>>>(EP, D.tabs[0], D.tabs[1], D.tabs[2], pairs_d, Evv); break;
        default: k_element_tables<0, 0, 0><<<eg, 256, 0, c->stream>>>(EP, D.tabs[0], D.tabs[1], D.tabs[2], pairs_d, Evv); break;
        }
        JFEM_HIP(hipGetLastError());
        JFEM_HIP(hipStreamSynchronize(c->stream));
    }
    if (has_div || has_vg) {
        const long per = (long)d * nlv * nlp;
        std::vector<double> eb((size_t)ntuples * per), eg(has_div && has_vg ? per : 0);
        for (long t = 0; t < ntuples; t++) {
            int cl[MAXD] = {0, 0, 0};
            long r = t;
            for (int a = d - 1; a >= 0; a--) { cl[a] = (int)(r % CV.ncls[a]); r /= CV.ncls[a]; }
            const Tab1D* tv[MAXD] = {&CV.tabs[0][cl[0]], &CV.tabs[1][cl[1]], d > 2 ? &CV.tabs[2][cl[2]] : nullptr};
            const Tab1D* tq[MAXD] = {&tabp[0][cl[0]], &tabp[1][cl[1]], d > 2 ? &tabp[2][cl[2]] : nullptr};
            double* et = eb.data() + t * per;
            if (has_div) div_value_table(d, gv.k, gp.k, Q, tv, tq, div_w, et);
            if (has_vg && !has_div) div_value_table(d, gv.k, gp.k, Q, tv, tq, vg_w, et, true);
            if (has_vg && has_div) {
                div_value_table(d, gv.k, gp.k, Q, tv, tq, vg_w, eg.data(), true);
                for (long i = 0; i < per; i++) et[i] += eg[i];
            }
        }
        EB.alloc(eb.size());
        // on the context stream: a blocking pageable copy pays a one-time runtime set-up (12 ms)
        JFEM_HIP(hipMemcpyAsync(EB, eb.data(), sizeof(double) * eb.size(), hipMemcpyHostToDevice, c->stream));
        JFEM_HIP(hipStreamSynchronize(c->stream));
    }
    ThParams P;
    std::memset(&P, 0, sizeof(P));
    P.kv = gv.k; P.kp = gp.k; P.nlocv = nlv; P.nlocp = nlp;
    P.nv = gv.ndofs(); P.np = gp.ndofs();
    for (int a = 0; a < MAXD; a++) {
        P.n[a] = a < d ? gv.n[a] : 1;
        P.NV[a] = a < d ? (int)gv.nodes_axis(a) : 1;
        P.NP[a] = a < d ? (int)gp.nodes_axis(a) : 1;
        P.ncls[a] = a < d ? CV.ncls[a] : 1;
        P.cls[a] = D.cls[a];
    }
    P.Evv = has_vv ? Evv.p : nullptr;
    P.EB = (has_div || has_vg) ? EB.p : nullptr;
    const long n = d * P.nv + P.np;
    JFEM_CHECK(n < (1l << 31), JFEM_ERR_UNSUPPORTED, "2^31 or more unknowns");
    auto* A = new jfem_mat();
    A->ctx = c;
    A->rows = A->cols = n;
    try {
        DevBuf<int> counts(n), bad(1);
        JFEM_HIP(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
        JFEM_HIP(hipMalloc(&A->rowptr, sizeof(int) * (n + 1)));
        const int blocks = grid_for(n, 256, 65536);
        auto launch = [&](bool fill) {
            if (d == 2) {
                if (fill) k_th_rows<2, true><<<blocks, 256, 0, c->stream>>>(P, n, counts, A->rowptr, A->colidx, A->vals, bad);
                else k_th_rows<2, false><<<blocks, 256, 0, c->stream>>>(P, n, counts, A->rowptr, A->colidx, A->vals, bad);
            } else {
                if (fill) k_th_rows<3, true><<<blocks, 256, 0, c->stream>>>(P, n, counts, A->rowptr, A->colidx, A->vals, bad);
                else k_th_rows<3, false><<<blocks, 256, 0, c->stream>>>(P, n, counts, A->rowptr, A->colidx, A->vals, bad);
            }
            JFEM_HIP(hipGetLastError());
        };
        launch(false);
        A->nnz = scan_counts(c, counts, n, A->rowptr);
        JFEM_HIP(hipMalloc(&A->colidx, sizeof(int) * (A->nnz + NNZ_PAD)));
        JFEM_HIP(hipMalloc(&A->vals, sizeof(double) * (A->nnz + NNZ_PAD)));
        launch(true);
        int bad_h = 0;
        JFEM_HIP(hipMemcpyAsync(&bad_h, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        JFEM_HIP(hipStreamSynchronize(c->stream));
        JFEM_CHECK(!bad_h, JFEM_ERR_STATE, "assembled rows differ from their reserved lengths");
        A->lanes = choose_lanes(n > 0 ? (double)A->nnz / n : 1.0);
        mat_build_blocks(A);
    } catch (...) {
        mat_free(A);
        throw;
    }
    return A;
}

// scatter a scalar vector (DOF order) into slice [off, off + n) of dst in node order, adding
__global__ void k_scatter_add_lex(long n, const int* __restrict__ rank, const double* __restrict__ src, long off,
                                  double* __restrict__ dst)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long j = off + rank[i];
        dst[j] = dst[j] + src[i];
    }
}
}  // namespace jfem

using namespace jfem;

extern "C" {

int jfem_space_create(jfem_ctx* ctx, int dim, const int* cells, const double* start, const double* end, int degree,
                      jfem_space** out)
{
    API_BEGIN
    JFEM_CHECK(ctx && cells && start && end && out, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(dim >= 1 && dim <= MAXD, JFEM_ERR_ARG, "Dimensions of coordinates or cell counts do not match");
    JFEM_CHECK(degree >= 1 && degree <= 4, JFEM_ERR_UNSUPPORTED, "polynomial degree must be 1..4");
    Grid g;
    g.d = dim;
    g.k = degree;
    for (int a = 0; a < dim; a++) {
        JFEM_CHECK(cells[a] >= 1, JFEM_ERR_ARG, "Range: lower index higher than upper");
        JFEM_CHECK(cells[a] <= (1 << 20), JFEM_ERR_ARG, "too many cells along an axis");
        g.n[a] = cells[a];
        g.start[a] = start[a];
        g.end[a] = end[a];
    }
    JFEM_CHECK(g.ndofs() < (1l << 31) - 1, JFEM_ERR_UNSUPPORTED, "2^31 or more degrees of freedom");
    auto* s = new jfem_space();
    s->ctx = ctx;
    s->grid = g;
    *out = s;
    API_END
}

int jfem_space_destroy(jfem_space* s)
{
    API_BEGIN
    delete s;
    API_END
}

long jfem_space_ndofs(const jfem_space* s) { return s ? s->grid.ndofs() : -1; }

int jfem_space_cell_dofs(jfem_space* s, int* out)
{
    API_BEGIN
    JFEM_CHECK(s && out, JFEM_ERR_ARG, "null argument");
    const long nc = s->grid.ncells();
    const int nloc = s->grid.nloc();
    int* d = nullptr;
    JFEM_HIP(hipMalloc(&d, sizeof(int) * nc * nloc));
    k_cell_dofs<<<grid_for(nc * nloc, 256, 16384), 256, 0, s->ctx->stream>>>(to_dev(s->grid), nc, nloc, d);
    hipError_t e = hipMemcpyAsync(out, d, sizeof(int) * nc * nloc, hipMemcpyDeviceToHost, s->ctx->stream);
    hipError_t e2 = hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(d);
    JFEM_HIP(e);
    JFEM_HIP(e2);
    API_END
}

int jfem_space_boundary_nodes(jfem_space* s, int* out, long* count)
{
    API_BEGIN
    JFEM_CHECK(s && out && count, JFEM_ERR_ARG, "null argument");
    const long n = s->grid.ndofs();
    unsigned char* m = nullptr;
    JFEM_HIP(hipMalloc(&m, n));
    space_boundary_mask(s, m);
    std::vector<unsigned char> h(n);
    hipError_t e = hipMemcpyAsync(h.data(), m, n, hipMemcpyDeviceToHost, s->ctx->stream);
    hipError_t e2 = hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(m);
    JFEM_HIP(e);
    JFEM_HIP(e2);
    long nb = 0;
    for (long i = 0; i < n; i++) if (h[i]) out[nb++] = (int)i;
    *count = nb;
    API_END
}

int jfem_space_node_coords(jfem_space* s, double* out)
{
    API_BEGIN
    JFEM_CHECK(s && out, JFEM_ERR_ARG, "null argument");
    const Grid& g = s->grid;
    const long n = g.ndofs();
    double* d = nullptr;
    double* se[2 * MAXD] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    JFEM_HIP(hipMalloc(&d, sizeof(double) * n * g.d));
    for (int a = 0; a < g.d; a++) {
        std::vector<double> s1, e1;
        cells_1d(g.start[a], g.end[a], g.n[a], s1, e1);
        JFEM_HIP(hipMalloc(&se[a], sizeof(double) * g.n[a]));
        JFEM_HIP(hipMalloc(&se[MAXD + a], sizeof(double) * g.n[a]));
        JFEM_HIP(hipMemcpyAsync(se[a], s1.data(), sizeof(double) * g.n[a], hipMemcpyHostToDevice, s->ctx->stream));
        JFEM_HIP(hipMemcpyAsync(se[MAXD + a], e1.data(), sizeof(double) * g.n[a], hipMemcpyHostToDevice,
                                s->ctx->stream));
        JFEM_HIP(hipStreamSynchronize(s->ctx->stream));
    }
    k_node_coords<<<grid_for(n, 256, 16384), 256, 0, s->ctx->stream>>>(to_dev(g), n, se[0], se[1], se[2], se[3], se[4],
                                                                      se[5], d);
    hipError_t e = hipMemcpyAsync(out, d, sizeof(double) * n * g.d, hipMemcpyDeviceToHost, s->ctx->stream);
    hipError_t e2 = hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(d);
    for (auto p : se) (void)hipFree(p);
    JFEM_HIP(e);
    JFEM_HIP(e2);
    API_END
}

int jfem_space_evaluate(jfem_space* s, const jfem_vec* coeffs, long npts, const double* points, double* values)
{
    API_BEGIN
    JFEM_CHECK(s && coeffs && (npts == 0 || (points && values)), JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(npts >= 0, JFEM_ERR_ARG, "negative point count");
    JFEM_CHECK(coeffs->n == s->grid.ndofs(), JFEM_ERR_ARG, "coefficient vector size differs from the space");
    if (npts == 0) return JFEM_OK;
    const int d = s->grid.d;
    DevBuf<double> p(npts * d), v(npts);
    JFEM_HIP(hipMemcpyAsync(p, points, sizeof(double) * npts * d, hipMemcpyHostToDevice, s->ctx->stream));
    space_evaluate_dev(s, coeffs->d, npts, p, v);
    JFEM_HIP(hipMemcpyAsync(values, v, sizeof(double) * npts, hipMemcpyDeviceToHost, s->ctx->stream));
    JFEM_HIP(hipStreamSynchronize(s->ctx->stream));
    API_END
}

int jfem_fe_l2_difference(jfem_space* s1, const jfem_vec* c1, jfem_space* s2, const jfem_vec* c2, long npts,
                          const double* points, double* result)
{
    API_BEGIN
    JFEM_CHECK(s1 && c1 && s2 && c2 && result && (npts == 0 || points), JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(npts >= 0, JFEM_ERR_ARG, "negative point count");
    JFEM_CHECK(s1->grid.d == s2->grid.d, JFEM_ERR_ARG, "the two spaces have different dimensions");
    JFEM_CHECK(c1->n == s1->grid.ndofs() && c2->n == s2->grid.ndofs(), JFEM_ERR_ARG,
               "coefficient vector size differs from its space");
    *result = 0;
    if (npts == 0) return JFEM_OK;
    const int d = s1->grid.d;
    DevBuf<double> p(npts * d), v1(npts), v2(npts);
    JFEM_HIP(hipMemcpyAsync(p, points, sizeof(double) * npts * d, hipMemcpyHostToDevice, s1->ctx->stream));
    space_evaluate_dev(s1, c1->d, npts, p, v1);
    space_evaluate_dev(s2, c2->d, npts, p, v2);
    // squaredDifferences = sum_i (f1(p_i) - f2(p_i))^2 (ConvergenceOrderEstimator.java:49-57), as the
    // dot product of the differences with themselves (the reference sums sequentially; the values
    // themselves are sums in nondeterministic order there)
    k_diff<<<grid_for(npts, 256, 16384), 256, 0, s1->ctx->stream>>>(npts, v1, v2, v1);
    JFEM_HIP(hipGetLastError());
    const double acc = vec_dot(s1->ctx, npts, v1, v1);
    *result = std::sqrt(acc);
    API_END
}

int jfem_assemble(jfem_space* s, int nint, const int* kinds, const double* weights, const double* cell_coeff,
                  int quad_rule, jfem_mat** out)
{
    API_BEGIN
    JFEM_CHECK(s && kinds && weights && out, JFEM_ERR_ARG, "null argument");
    *out = assemble_matrix(s, nint, kinds, weights, cell_coeff, quad_rule);
    API_END
}

int jfem_assemble_rhs_sines(jfem_space* s, int nterms, const int* kw, const double* amp, int quad_rule, jfem_vec* rhs)
{
    API_BEGIN
    JFEM_CHECK(s && kw && amp && rhs, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(rhs->n == s->grid.ndofs(), JFEM_ERR_ARG, "Incompatible sizes");
    assemble_rhs_sines(s, nterms, kw, amp, quad_rule, rhs->d, nullptr);
    API_END
}

int jfem_assemble_rhs_trig(jfem_space* s, int nterms, const int* kw, const int* cosine, const double* amp,
                           int quad_rule, jfem_vec* rhs)
{
    API_BEGIN
    JFEM_CHECK(s && kw && amp && rhs, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(rhs->n == s->grid.ndofs(), JFEM_ERR_ARG, "Incompatible sizes");
    assemble_rhs_sines(s, nterms, kw, amp, quad_rule, rhs->d, cosine);
    API_END
}

int jfem_space_lex_rank(jfem_space* s, int* out)
{
    API_BEGIN
    JFEM_CHECK(s && out, JFEM_ERR_ARG, "null argument");
    const long n = s->grid.ndofs();
    DevBuf<int> r(n);
    space_lex_rank(s, r);
    // the context stream is non-blocking: copy on it, then wait
    JFEM_HIP(hipMemcpyAsync(out, r, sizeof(int) * n, hipMemcpyDeviceToHost, s->ctx->stream));
    JFEM_HIP(hipStreamSynchronize(s->ctx->stream));
    API_END
}

int jfem_th_assemble(jfem_space* velocity, jfem_space* pressure, int nint, const int* kinds, const double* weights,
                     int quad_rule, jfem_mat** out)
{
    API_BEGIN
    JFEM_CHECK(velocity && pressure && out && (nint == 0 || (kinds && weights)), JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(velocity->ctx == pressure->ctx, JFEM_ERR_ARG, "spaces on different contexts");
    bool has_vv = false, has_div = false, has_vg = false;
    double wv = 0, wd = 0, wg = 0;
    for (int i = 0; i < nint; i++) {
        if (kinds[i] == JFEM_MIXED_VELOCITY_GRAD_GRAD) {
            JFEM_CHECK(!has_vv, JFEM_ERR_UNSUPPORTED, "one velocity GRAD_GRAD integral at most");
            has_vv = true;
            wv = weights[i];
        } else if (kinds[i] == JFEM_MIXED_DIV_VALUE) {
            JFEM_CHECK(!has_div, JFEM_ERR_UNSUPPORTED, "one DIV_VALUE integral at most");
            has_div = true;
            wd = weights[i];
        } else if (kinds[i] == JFEM_MIXED_VALUE_GRAD) {
            JFEM_CHECK(!has_vg, JFEM_ERR_UNSUPPORTED, "one VALUE_GRAD integral at most");
            has_vg = true;
            wg = weights[i];
        } else {
            JFEM_CHECK(false, JFEM_ERR_UNSUPPORTED, "Integral type not supported");
        }
    }
    *out = th_assemble(velocity, pressure, has_vv, wv, has_div, wd, quad_rule, has_vg, wg);
    API_END
}

int jfem_th_assemble_rhs_trig(jfem_space* velocity, jfem_space* pressure, int component, int nterms, const int* kw,
                              const int* cosine, const double* amp, int quad_rule, jfem_vec* rhs)
{
    API_BEGIN
    JFEM_CHECK(velocity && pressure && kw && amp && rhs, JFEM_ERR_ARG, "null argument");
    const int d = velocity->grid.d;
    JFEM_CHECK(component >= 0 && component < d, JFEM_ERR_ARG, "velocity component out of range");
    const long nv = velocity->grid.ndofs();
    JFEM_CHECK(rhs->n == d * nv + pressure->grid.ndofs(), JFEM_ERR_ARG, "Incompatible sizes");
    // TPVectorRightHandSideIntegral evaluates value(x) of the whole shape function (see th_assemble)
    JFEM_CHECK(quad_rule >= 1 && quad_rule <= 5, JFEM_ERR_UNSUPPORTED,
               "mixed integrals need a Gauss rule (Gauss-Lobatto points lie on cell faces)");
    DevBuf<double> tmp(nv);
    DevBuf<int> rank(nv);
    assemble_rhs_sines(velocity, nterms, kw, amp, quad_rule, tmp, cosine);
    space_lex_rank(velocity, rank);
    k_scatter_add_lex<<<grid_for(nv, 256, 16384), 256, 0, velocity->ctx->stream>>>(nv, rank, tmp, component * nv,
                                                                                     rhs->d);
    JFEM_HIP(hipGetLastError());
    JFEM_HIP(hipStreamSynchronize(velocity->ctx->stream));
    API_END
}

}  // extern "C"
