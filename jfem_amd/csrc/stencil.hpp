// Matrix-free operator of a multigrid level (stencil-aware smoothing and residuals).
//
// A level system of the hierarchy is the Dirichlet-eliminated assembly of
// TPCellIntegral GRAD_GRAD / VALUE_VALUE / H1 on a structured Qk grid with a
// piecewise-constant cell coefficient a_c (MGPreconditionerSpace.createSystem,
// M/multigrid/MGPreconditionerSpace.java:28-57).  Tensor-product Gauss quadrature
// factorises over the axes, so the element matrix of cell c is
//     a_c [ w_gg sum_a K_a (x) M_b (x) ... + w_vv M_0 (x) M_1 (x) ... ]
// with K_a, M_a the 1D stiffness and mass matrices of one cell of axis a.  After
// the elimination, a boundary row is the identity and boundary columns are zero.
//
// One workgroup owns a tile of cells (the nodes they own in the reference's
// first-appearance numbering).  It
//   1. gathers x of the nodes of the tile's cells and of one more cell layer on the
//      high side of every axis (boundary nodes as 0) from DOF order into an LDS
//      lattice tile, and the coefficients of those cells;
//   2. applies every such cell's element matrix by 1D sum factorisation, one cell
//      per thread, and keeps the (k+1)^d local results per cell in LDS;
//   3. for every owned node sums the results of its (up to 2^d) cells in a fixed
//      order (no atomics: deterministic run to run) and hands (A x)_row to the same
//      per-row epilogues as the CSR-stream kernel (residual, Jacobi, Chebyshev...).
// No matrix is read: per node the kernel streams x, the epilogue operands and the
// outputs, against 12 bytes per stored entry (about 63 per row in 3D Q2) for CSR.
#pragma once

#include "common.hpp"

namespace jfem {

constexpr int MF_BS = 256;   // threads per workgroup

// per-level parameters
struct MfParams {
    GridDev g;
    int ntile[MAXD];         // tiles per axis
    double K1[MAXD][9];      // 1D stiffness of one cell per axis, (k+1) x (k+1) row-major
    double M1[MAXD][9];      // 1D mass
    double wgg, wvv;         // weights of the gradient and value terms
};

// cells per axis of a tile (the computed box has one more per axis)
template <int D, int K>
struct MfShape;
template <> struct MfShape<3, 2> { static constexpr int T0 = 4, T1 = 4, T2 = 4; };
template <> struct MfShape<3, 1> { static constexpr int T0 = 4, T1 = 8, T2 = 8; };
template <> struct MfShape<2, 2> { static constexpr int T0 = 8, T1 = 16, T2 = 1; };
template <> struct MfShape<2, 1> { static constexpr int T0 = 16, T1 = 16, T2 = 1; };

// the reference's DOF index of lattice point p (dof_of_lattice with d and k fixed)
template <int D, int K>
__device__ inline long mf_dof(const int* n, const int* p)
{
    long tail[D];
    long t = 1;
#pragma unroll
    for (int a = D - 1; a >= 0; a--) { tail[a] = t; t *= (long)K * n[a] + 1; }
    long pre = 0, mprod = 1, rank = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const int c = p[a] == 0 ? 0 : (p[a] - 1) / K;
        const int l = p[a] - c * K;
        pre += mprod * (c > 0 ? (long)c * K + 1 : 0) * tail[a];
        rank += (long)(l - (c > 0 ? 1 : 0)) * mprod;
        mprod *= c > 0 ? K : K + 1;
    }
    return pre + rank;
}

// y = (w_gg sum_a K_a (x) M_b + w_vv (x)_a M_a) x on one cell by sum factorisation.  xs:
// the cell's node (0, 0[, 0]) in the LDS tile, strides s0, s1 (axis 2 contiguous); y local,
// axis 0 fastest: y[l0 + N l1 + N^2 l2]
template <typename T, int D, int K>
__device__ inline void mf_cell(const MfParams& P, const T* xs, int s0, int s1, T* y)
{
    constexpr int N = K + 1;
    const T wgg = (T)P.wgg, wvv = (T)P.wvv;
    if constexpr (D == 2) {
        // u(j0, i1) = sum_j1 {M,K}_1(i1, j1) x(j0, j1)
        T uM[N][N], uK[N][N];
#pragma unroll
        for (int j0 = 0; j0 < N; j0++)
#pragma unroll
            for (int i1 = 0; i1 < N; i1++) {
                T m = 0, k = 0;
#pragma unroll
                for (int j1 = 0; j1 < N; j1++) {
                    const T xv = xs[j0 * s0 + j1 * s1];
                    m += (T)P.M1[1][i1 * N + j1] * xv;
                    k += (T)P.K1[1][i1 * N + j1] * xv;
                }
                uM[j0][i1] = m;
                uK[j0][i1] = k;
            }
#pragma unroll
        for (int i0 = 0; i0 < N; i0++)
#pragma unroll
            for (int i1 = 0; i1 < N; i1++) {
                T acc = 0;
#pragma unroll
                for (int j0 = 0; j0 < N; j0++)
                    acc += (T)P.K1[0][i0 * N + j0] * (wgg * uM[j0][i1]) +
                           (T)P.M1[0][i0 * N + j0] * (wgg * uK[j0][i1] + wvv * uM[j0][i1]);
                y[i0 + N * i1] = acc;
            }
    } else {
        // per slab j0 of the cell: A = w_gg (M_1 (x) M_2) x_j0, B = w_gg (K_1 (x) M_2 + M_1 (x) K_2) x_j0
        // + w_vv (M_1 (x) M_2) x_j0; then y(i0, .) += K_0(i0, j0) A + M_0(i0, j0) B
#pragma unroll
        for (int l = 0; l < N * N * N; l++) y[l] = 0;
#pragma unroll
        for (int j0 = 0; j0 < N; j0++) {
            T uM[N][N], uK[N][N];   // [j1][i2]
#pragma unroll
            for (int j1 = 0; j1 < N; j1++)
#pragma unroll
                for (int i2 = 0; i2 < N; i2++) {
                    T m = 0, k = 0;
#pragma unroll
                    for (int j2 = 0; j2 < N; j2++) {
                        const T xv = xs[j0 * s0 + j1 * s1 + j2];
                        m += (T)P.M1[2][i2 * N + j2] * xv;
                        k += (T)P.K1[2][i2 * N + j2] * xv;
                    }
                    uM[j1][i2] = m;
                    uK[j1][i2] = k;
                }
#pragma unroll
            for (int i1 = 0; i1 < N; i1++)
#pragma unroll
                for (int i2 = 0; i2 < N; i2++) {
                    T mm = 0, km = 0, mk = 0;
#pragma unroll
                    for (int j1 = 0; j1 < N; j1++) {
                        mm += (T)P.M1[1][i1 * N + j1] * uM[j1][i2];
                        km += (T)P.K1[1][i1 * N + j1] * uM[j1][i2];
                        mk += (T)P.M1[1][i1 * N + j1] * uK[j1][i2];
                    }
                    const T A = wgg * mm, B = wgg * (km + mk) + wvv * mm;
#pragma unroll
                    for (int i0 = 0; i0 < N; i0++)
                        y[i0 + N * i1 + N * N * i2] += (T)P.K1[0][i0 * N + j0] * A + (T)P.M1[0][i0 * N + j0] * B;
                }
        }
    }
}

// Epi: the row epilogues of stream.hpp (Pre load(row), operator()(row, ax, pre)); coef: one
// value per cell of the level (reference cell order, last axis fastest), nullptr = 1
template <typename T, int D, int K, typename Epi>
__global__ __launch_bounds__(MF_BS) void k_mf(MfParams P, const T* __restrict__ x, const T* __restrict__ coef, Epi epi)
{
    static_assert(!Epi::reduce, "matrix-free epilogues do not reduce");
    using S = MfShape<D, K>;
    constexpr int N = K + 1;
    constexpr int NL = D == 3 ? N * N * N : N * N;
    constexpr int TC[3] = {S::T0, S::T1, D == 3 ? S::T2 : 1};
    constexpr int CC1 = S::T1 + 1, CC2 = D == 3 ? S::T2 + 1 : 1;
    constexpr int CC0 = S::T0 + 1;
    constexpr int NB1 = CC1 * K + 1, NB2 = D == 3 ? CC2 * K + 1 : 1;
    constexpr int NB0 = CC0 * K + 1;
    __shared__ T xs[NB0 * NB1 * NB2];
    __shared__ T cr[CC0 * CC1 * CC2 * NL];
    __shared__ T cs[CC0 * CC1 * CC2];
    const int tid = threadIdx.x;
    const int* n = P.g.n;
    // tile coordinates, last axis fastest (consecutive workgroups own consecutive DOF runs)
    int clo[3] = {0, 0, 0}, chi[3] = {1, 1, 1}, ncc[3] = {1, 1, 1}, nb[3] = {1, 1, 1}, org[3] = {0, 0, 0};
    {
        long b = blockIdx.x;
#pragma unroll
        for (int a = D - 1; a >= 0; a--) {
            const int t = (int)(b % P.ntile[a]);
            b /= P.ntile[a];
            clo[a] = t * TC[a];
            chi[a] = min(clo[a] + TC[a], n[a]);
            ncc[a] = min(chi[a] + 1, n[a]) - clo[a];
            nb[a] = ncc[a] * K + 1;
            org[a] = clo[a] * K;
        }
    }
    // 1. node values (boundary nodes as 0: their columns are eliminated) and cell coefficients
    const int nn = nb[0] * nb[1] * nb[2];
    for (int i = tid; i < nn; i += MF_BS) {
        const int q2 = i % nb[2], r = i / nb[2], q1 = r % nb[1], q0 = r / nb[1];
        const int p[3] = {org[0] + q0, org[1] + q1, org[2] + q2};
        bool bnd = false;
#pragma unroll
        for (int a = 0; a < D; a++) bnd |= p[a] == 0 || p[a] == K * n[a];
        xs[(q0 * NB1 + q1) * NB2 + q2] = bnd ? (T)0 : x[mf_dof<D, K>(n, p)];
    }
    const int nc = ncc[0] * ncc[1] * ncc[2];
    for (int i = tid; i < nc; i += MF_BS) {
        const int e2 = i % ncc[2], r = i / ncc[2], e1 = r % ncc[1], e0 = r / ncc[1];
        long cl = clo[0] + e0;
        cl = cl * n[1] + clo[1] + e1;
        if (D == 3) cl = cl * n[2] + clo[2] + e2;
        cs[(e0 * CC1 + e1) * CC2 + e2] = coef ? coef[cl] : (T)1;
    }
    __syncthreads();
    // 2. element matrices by sum factorisation, one cell per thread
    for (int i = tid; i < nc; i += MF_BS) {
        const int e2 = i % ncc[2], r = i / ncc[2], e1 = r % ncc[1], e0 = r / ncc[1];
        const int ce = (e0 * CC1 + e1) * CC2 + e2;
        T y[NL];
        mf_cell<T, D, K>(P, xs + ((e0 * K) * NB1 + e1 * K) * NB2 + (D == 3 ? e2 * K : 0), NB1 * NB2, NB2, y);
        const T a = cs[ce];
#pragma unroll
        for (int l = 0; l < NL; l++) cr[ce * NL + l] = a * y[l];
    }
    __syncthreads();
    // 3. owned nodes: lattice points whose owner cell lies in the tile
    int olo[3] = {0, 0, 0}, ono[3] = {1, 1, 1};
#pragma unroll
    for (int a = 0; a < D; a++) {
        olo[a] = clo[a] == 0 ? 0 : clo[a] * K + 1;
        ono[a] = chi[a] * K - olo[a] + 1;
    }
    const int no = ono[0] * ono[1] * ono[2];
    for (int i = tid; i < no; i += MF_BS) {
        const int o2 = i % ono[2], r = i / ono[2], o1 = r % ono[1], o0 = r / ono[1];
        const int p[3] = {olo[0] + o0, olo[1] + o1, olo[2] + o2};
        const long row = mf_dof<D, K>(n, p);
        const auto pre = epi.load(row);
        bool bnd = false;
#pragma unroll
        for (int a = 0; a < D; a++) bnd |= p[a] == 0 || p[a] == K * n[a];
        T ax;
        if (bnd) {
            ax = x[row];   // identity row of the eliminated system
        } else {
            // the cells containing p, ascending; all lie in the computed box
            int q[3], elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
#pragma unroll
            for (int a = 0; a < 3; a++) {
                q[a] = p[a] - org[a];
                if (a < D) {
                    ehi[a] = q[a] / K;
                    elo[a] = q[a] % K == 0 ? ehi[a] - 1 : ehi[a];
                }
            }
            T acc = 0;
            for (int e0 = elo[0]; e0 <= ehi[0]; e0++)
                for (int e1 = elo[1]; e1 <= ehi[1]; e1++)
                    for (int e2 = elo[2]; e2 <= ehi[2]; e2++) {
                        const int l = (q[0] - e0 * K) + N * (q[1] - e1 * K) + (D == 3 ? N * N * (q[2] - e2 * K) : 0);
                        acc += cr[((e0 * CC1 + e1) * CC2 + e2) * NL + l];
                    }
            ax = acc;
        }
        epi(row, ax, pre);
    }
}

// y-row epilogue launch over a level's tiles; returns false for (d, k) without a kernel
template <typename T, typename Epi>
inline bool launch_mf(hipStream_t st, const MfParams& P, const T* x, const T* coef, const Epi& epi)
{
    long tiles = 1;
    for (int a = 0; a < P.g.d; a++) tiles *= P.ntile[a];
    if (tiles == 0) return true;
    JFEM_CHECK(tiles < (1l << 31), JFEM_ERR_UNSUPPORTED, "too many tiles");
    switch (P.g.d * 10 + P.g.k) {
    case 32: k_mf<T, 3, 2, Epi><<<(unsigned)tiles, MF_BS, 0, st>>>(P, x, coef, epi); break;
    case 31: k_mf<T, 3, 1, Epi><<<(unsigned)tiles, MF_BS, 0, st>>>(P, x, coef, epi); break;
    case 22: k_mf<T, 2, 2, Epi><<<(unsigned)tiles, MF_BS, 0, st>>>(P, x, coef, epi); break;
    case 21: k_mf<T, 2, 1, Epi><<<(unsigned)tiles, MF_BS, 0, st>>>(P, x, coef, epi); break;
    default: return false;
    }
    JFEM_HIP(hipGetLastError());
    return true;
}

// tiles per axis for a grid
inline void mf_tiles(const Grid& g, int* ntile)
{
    int T[3] = {1, 1, 1};
    if (g.d == 3 && g.k == 2) { T[0] = MfShape<3, 2>::T0; T[1] = MfShape<3, 2>::T1; T[2] = MfShape<3, 2>::T2; }
    if (g.d == 3 && g.k == 1) { T[0] = MfShape<3, 1>::T0; T[1] = MfShape<3, 1>::T1; T[2] = MfShape<3, 1>::T2; }
    if (g.d == 2 && g.k == 2) { T[0] = MfShape<2, 2>::T0; T[1] = MfShape<2, 2>::T1; }
    if (g.d == 2 && g.k == 1) { T[0] = MfShape<2, 1>::T0; T[1] = MfShape<2, 1>::T1; }
    for (int a = 0; a < MAXD; a++) ntile[a] = a < g.d ? (g.n[a] + T[a] - 1) / T[a] : 1;
}

inline bool mf_supported(const Grid& g) { return (g.d == 2 || g.d == 3) && (g.k == 1 || g.k == 2); }

}  // namespace jfem
