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
//   0. issues the loads of the epilogue operands of its owned rows;
//   1. gathers x of the nodes of the tile's cells and of one more cell layer on the
//      high side of every axis (boundary nodes as 0) from DOF order into an LDS
//      lattice tile, and the coefficients of those cells;
//   2. applies every such cell's element matrix by 1D sum factorisation, k+1 threads
//      per cell (split by the output index along the last axis, no work repeated),
//      and keeps the (k+1)^d local results per cell in LDS;
//   3. for every owned node sums the results of its (up to 2^d) cells in a fixed
//      order (no atomics: deterministic run to run) and hands (A x)_row to the same
//      per-row epilogues as the CSR-stream kernel (residual, Jacobi, Chebyshev...).
// No matrix is read: per node the kernel streams x, the epilogue operands and the
// outputs, against 12 bytes per stored entry (about 63 per row in 3D Q2) for CSR.
#pragma once

#include "common.hpp"

#include <type_traits>
#include <vector>

namespace jfem {

// per-level parameters
struct MfParams {
    GridDev g;
    int ntile[MAXD];         // tiles per axis
    double K1[MAXD][9];      // 1D stiffness of one cell per axis, (k+1) x (k+1) row-major
    double M1[MAXD][9];      // 1D mass
    double wgg, wvv;         // weights of the gradient and value terms
    // table of the inner tiles (tiles whose node box starts at cell 2 or later and ends before the
    // last lattice point on every axis: full tiles without boundary nodes), built by mf_tables:
    // cand[8 q + c] = the offset in the tile's cell results of candidate cell c of node-box
    // position q (0xffff: none), the same for every inner tile
    const unsigned short* cand;
};

// cells per axis of a tile (the computed box has one more per axis) and threads per
// workgroup: k+1 threads share a computed cell, so BS >= (k+1) * computed cells
template <int D, int K>
struct MfShape;
template <> struct MfShape<3, 2> { static constexpr int T0 = 4, T1 = 4, T2 = 4, BS = 384; };
template <> struct MfShape<3, 1> { static constexpr int T0 = 4, T1 = 4, T2 = 8, BS = 512; };
template <> struct MfShape<2, 2> { static constexpr int T0 = 8, T1 = 16, T2 = 1, BS = 512; };
template <> struct MfShape<2, 1> { static constexpr int T0 = 16, T1 = 12, T2 = 1, BS = 512; };

// consecutive tiles per XCD (as ST_XCD_RUN for the CSR-stream blocks): neighbouring tiles
// share a cell layer of x
constexpr int MF_XCD_RUN = 16;

// the reference's DOF index of lattice point p (dof_of_lattice with d and k fixed; every
// index is below 2^31, as the CSR of the level requires)
template <int D, int K>
__device__ inline int mf_dof(const int* n, const int* p)
{
    int tail[D];
    int t = 1;
#pragma unroll
    for (int a = D - 1; a >= 0; a--) { tail[a] = t; t *= K * n[a] + 1; }
    int pre = 0, mprod = 1, rank = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const int c = p[a] == 0 ? 0 : (p[a] - 1) / K;
        const int l = p[a] - c * K;
        pre += mprod * (c > 0 ? c * K + 1 : 0) * tail[a];
        rank += (l - (c > 0 ? 1 : 0)) * mprod;
        mprod *= c > 0 ? K : K + 1;
    }
    return pre + rank;
}

// row r of a (k+1) x (k+1) table with r varying per lane, as registers
template <int N>
__device__ inline void mf_row(const double* tab, int r, double* out)
{
#pragma unroll
    for (int j = 0; j < N; j++) {
        double v = tab[j];
#pragma unroll
        for (int i = 1; i < N; i++) v = r == i ? tab[i * N + j] : v;
        out[j] = v;
    }
}

// one computed cell, the N outputs (2D) or N x N outputs (3D) with last-axis index s:
// y = a_c (w_gg sum_a K_a (x) M_b + w_vv (x)_a M_a) x by 1D sum factorisation.  xs: the
// cell's node (0, 0[, 0]) in the LDS tile, strides s0, s1 (the last axis contiguous)
template <typename T, int D, int K>
__device__ inline void mf_cell_part(const MfParams& P, const T* xs, int s0, int s1, int s, T* y)
{
    // the library is built without contraction (the assembly follows the reference's operation
    // order); this operator is only held to rounding against the assembled matrix, so its
    // products and sums contract to fused multiply-adds
#pragma clang fp contract(fast)
    constexpr int N = K + 1;
    const T wgg = (T)P.wgg, wvv = (T)P.wvv;
    double mr[N], kr[N];   // row s of the last axis's tables
    mf_row<N>(P.M1[D - 1], s, mr);
    mf_row<N>(P.K1[D - 1], s, kr);
    if constexpr (D == 2) {
#pragma unroll
        for (int i0 = 0; i0 < N; i0++) y[i0] = 0;
#pragma unroll
        for (int j0 = 0; j0 < N; j0++) {
            T uM = 0, uK = 0;
#pragma unroll
            for (int j1 = 0; j1 < N; j1++) {
                const T xv = xs[j0 * s0 + j1];
                uM += (T)mr[j1] * xv;
                uK += (T)kr[j1] * xv;
            }
            const T A = wgg * uM, B = wgg * uK + wvv * uM;
#pragma unroll
            for (int i0 = 0; i0 < N; i0++) y[i0] += (T)P.K1[0][i0 * N + j0] * A + (T)P.M1[0][i0 * N + j0] * B;
        }
    } else {
#pragma unroll
        for (int l = 0; l < N * N; l++) y[l] = 0;   // y[i0 + N i1]
#pragma unroll
        for (int j0 = 0; j0 < N; j0++) {
            T uM[N], uK[N];
#pragma unroll
            for (int j1 = 0; j1 < N; j1++) {
                T m = 0, k = 0;
#pragma unroll
                for (int j2 = 0; j2 < N; j2++) {
                    const T xv = xs[j0 * s0 + j1 * s1 + j2];
                    m += (T)mr[j2] * xv;
                    k += (T)kr[j2] * xv;
                }
                uM[j1] = m;
                uK[j1] = k;
            }
#pragma unroll
            for (int i1 = 0; i1 < N; i1++) {
                T mm = 0, km = 0, mk = 0;
#pragma unroll
                for (int j1 = 0; j1 < N; j1++) {
                    mm += (T)P.M1[1][i1 * N + j1] * uM[j1];
                    km += (T)P.K1[1][i1 * N + j1] * uM[j1];
                    mk += (T)P.M1[1][i1 * N + j1] * uK[j1];
                }
                const T A = wgg * mm, B = wgg * (km + mk) + wvv * mm;
#pragma unroll
                for (int i0 = 0; i0 < N; i0++)
                    y[i0 + N * i1] += (T)P.K1[0][i0 * N + j0] * A + (T)P.M1[0][i0 * N + j0] * B;
            }
        }
    }
}

// epilogues whose operand x is the product's own (tile_x): k_mf loads the rest of the row operands
// (load_nox) and takes x at the row from its node tile instead of a second global read
template <typename E, typename = void>
struct EpiTileX : std::false_type {};
template <typename E>
struct EpiTileX<E, std::void_t<decltype(E::tile_x)>> : std::bool_constant<E::tile_x> {};

template <typename E>
__device__ inline typename E::Pre epi_load(const E& e, long row)
{
    if constexpr (EpiTileX<E>::value) return e.load_nox(row);
    else return e.load(row);
}

// Epi: the row epilogues of stream.hpp (Pre load(row), operator()(row, ax, pre)); coef: one
// value per cell of the level (reference cell order, last axis fastest), nullptr = 1
// waves per SIMD the kernel is built for: six in FP64 (four workgroups per CU), seven in FP32
template <typename T, typename Epi>
constexpr int mf_waves() { return sizeof(T) == 8 ? 6 : 7; }

template <typename T, int D, int K, typename Epi>
__global__ __launch_bounds__((MfShape<D, K>::BS), (mf_waves<T, Epi>())) void k_mf(MfParams P, const T* __restrict__ x,
                                                          const T* __restrict__ coef, Epi epi)
{
    static_assert(!Epi::reduce, "matrix-free epilogues do not reduce");
    using S = MfShape<D, K>;
    constexpr int BS = S::BS;
    constexpr int N = K + 1;
    constexpr int NL = D == 3 ? N * N * N : N * N;
    constexpr int TC[3] = {S::T0, S::T1, D == 3 ? S::T2 : 1};
    constexpr int CC0 = S::T0 + 1, CC1 = S::T1 + 1, CC2 = D == 3 ? S::T2 + 1 : 1;      // computed cells
    constexpr int NB0 = CC0 * K + 1, NB1 = CC1 * K + 1, NB2 = D == 3 ? CC2 * K + 1 : 1;   // their nodes
    constexpr int OB0 = S::T0 * K + 1, OB1 = S::T1 * K + 1, OB2 = D == 3 ? S::T2 * K + 1 : 1;   // owned nodes
    constexpr int NCC = CC0 * CC1 * CC2, NNB = NB0 * NB1 * NB2, NOB = OB0 * OB1 * OB2;
    constexpr int NOWN = (NOB + BS - 1) / BS;
    static_assert(BS >= N * NCC, "k+1 threads per computed cell");
    __shared__ T xs[NNB];
    __shared__ T cr[NCC * NL];
    __shared__ T cs[NCC];
    const int tid = threadIdx.x;
    const int* n = P.g.n;
    int bid = blockIdx.x;
    {
        constexpr int super = 8 * MF_XCD_RUN;
        const int full = (int)(gridDim.x / super) * super;
        if (bid < full) {
            const int m = bid >> 3;
            bid = (m / MF_XCD_RUN) * super + (bid & 7) * MF_XCD_RUN + (m % MF_XCD_RUN);
        }
    }
    // tile coordinates, last axis fastest (consecutive tiles own consecutive DOF runs)
    int clo[3] = {0, 0, 0}, chi[3] = {1, 1, 1}, ncc[3] = {1, 1, 1}, nb[3] = {1, 1, 1}, org[3] = {0, 0, 0};
    int olo[3] = {0, 0, 0}, ono[3] = {1, 1, 1};
    {
        int b = bid;
#pragma unroll
        for (int a = D - 1; a >= 0; a--) {
            const int t = b % P.ntile[a];
            b /= P.ntile[a];
            clo[a] = t * TC[a];
            chi[a] = min(clo[a] + TC[a], n[a]);
            ncc[a] = min(chi[a] + 1, n[a]) - clo[a];
            nb[a] = ncc[a] * K + 1;
            org[a] = clo[a] * K;
            olo[a] = clo[a] == 0 ? 0 : clo[a] * K + 1;   // owned lattice points: owner cell in the tile
            ono[a] = chi[a] * K - olo[a] + 1;
        }
    }
    // DOF index of box node q: for tiles whose node box starts past the first cell layer on every
    // axis (every node's owner cell has index >= 1 there), the reference's numbering is separable,
    // DOF(org + q) = g_0(q_0) + g_1(q_1) + g_2(q_2), with per-axis tables built once per tile;
    // other tiles use the closed form per node
    __shared__ int gtab[NB0 + NB1 + NB2];
    bool sep = true, inner = P.cand != nullptr;
#pragma unroll
    for (int a = 0; a < D; a++) {
        sep = sep && clo[a] >= 2;
        inner = inner && clo[a] >= 2 && clo[a] + TC[a] + 1 < n[a];
    }
    // inner tiles: the candidate cells of the owned nodes from cand (no per-node index arithmetic)
    if (sep) {
        const int base[3] = {K + 1, K + 1, D == 3 ? K + 1 : 0};
        const int b0 = mf_dof<D, K>(n, base);
        for (int i = tid; i < NB0 + NB1 + NB2; i += BS) {
            const int a = i < NB0 ? 0 : (i < NB0 + NB1 ? 1 : 2);
            const int q = i - (a == 0 ? 0 : (a == 1 ? NB0 : NB0 + NB1));
            if (a < D && q < nb[a]) {
                const int pp[3] = {a == 0 ? org[0] + q : base[0], a == 1 ? org[1] + q : base[1],
                                   a == 2 ? org[2] + q : base[2]};
                gtab[i] = mf_dof<D, K>(n, pp) - (a == 0 ? 0 : b0);
            }
        }
        __syncthreads();
    }
    auto dof_of = [&](int q0, int q1, int q2) -> int {
        if (sep) return gtab[q0] + gtab[NB0 + q1] + (D == 3 ? gtab[NB0 + NB1 + q2] : 0);
        const int pp[3] = {org[0] + q0, org[1] + q1, org[2] + q2};
        return mf_dof<D, K>(n, pp);
    };
    // 0. the owned rows of this thread and their epilogue operands, issued first; operands of more
    // than 16 bytes per row (the FP64 Chebyshev step's four) are loaded in phase 3, where they
    // are used: held from here they would not fit the registers of six waves per SIMD
    constexpr bool PRE_EARLY = sizeof(typename Epi::Pre) <= 16;
    int row[NOWN];
    bool own[NOWN];
    typename Epi::Pre pre[NOWN];
    // inner tiles: the owned node's candidate offsets, loaded with the row in FP64 (in FP32 the
    // registers of the seven workgroups per CU do not hold them; loaded where they are used)
    constexpr bool CAND_EARLY = sizeof(T) == 8;
    uint4 candw[NOWN];
#pragma unroll
    for (int u = 0; u < NOWN; u++) {
        const int i = tid + u * BS;
        const int o2 = i % OB2, o1 = (i / OB2) % OB1, o0 = i / (OB1 * OB2);
        own[u] = i < NOB && o0 < ono[0] && o1 < ono[1] && o2 < ono[2];
        row[u] = 0;
        if (own[u]) {
            row[u] = dof_of(olo[0] + o0 - org[0], olo[1] + o1 - org[1], olo[2] + o2 - org[2]);
            if (PRE_EARLY) pre[u] = epi_load(epi, row[u]);
            if (inner && CAND_EARLY)
                candw[u] = *reinterpret_cast<const uint4*>(
                    P.cand + 8 * (((o0 + 1) * NB1 + o1 + 1) * NB2 + (D == 3 ? o2 + 1 : 0)));
        }
    }
    // 1. node values (boundary nodes as 0: their columns are eliminated) and cell coefficients;
    // all loads of a thread are issued before any LDS store, so that they are in flight together
    constexpr int NPER = (NNB + BS - 1) / BS, CPER = (NCC + BS - 1) / BS;
    T xv[NPER], cv[CPER];
#pragma unroll
    for (int u = 0; u < NPER; u++) {
        const int i = tid + u * BS;
        const int q2 = i % NB2, q1 = (i / NB2) % NB1, q0 = i / (NB1 * NB2);
        xv[u] = 0;
        if (inner) {   // a full box without boundary nodes
            if (i < NNB) xv[u] = x[dof_of(q0, q1, q2)];
            continue;
        }
        if (i < NNB && q0 < nb[0] && q1 < nb[1] && q2 < nb[2]) {
            const int p[3] = {org[0] + q0, org[1] + q1, org[2] + q2};
            bool bnd = false;
#pragma unroll
            for (int a = 0; a < D; a++) bnd |= p[a] == 0 || p[a] == K * n[a];
            if (!bnd) xv[u] = x[dof_of(q0, q1, q2)];
        }
    }
#pragma unroll
    for (int u = 0; u < CPER; u++) {
        const int i = tid + u * BS;
        const int e2 = i % CC2, e1 = (i / CC2) % CC1, e0 = i / (CC1 * CC2);
        cv[u] = 1;
        if (coef && i < NCC && e0 < ncc[0] && e1 < ncc[1] && e2 < ncc[2]) {
            int cl = (clo[0] + e0) * n[1] + clo[1] + e1;
            if (D == 3) cl = cl * n[2] + clo[2] + e2;
            cv[u] = coef[cl];
        }
    }
#pragma unroll
    for (int u = 0; u < NPER; u++)
        if (tid + u * BS < NNB) xs[tid + u * BS] = xv[u];
#pragma unroll
    for (int u = 0; u < CPER; u++)
        if (tid + u * BS < NCC) cs[tid + u * BS] = cv[u];
    __syncthreads();
    // 2. element matrices by sum factorisation: k+1 threads per computed cell, split by the
    // output index along the last axis
    {
        const int ce = tid / N, s = tid % N;
        const int e2 = ce % CC2, e1 = (ce / CC2) % CC1, e0 = ce / (CC1 * CC2);
        if (ce < NCC && e0 < ncc[0] && e1 < ncc[1] && e2 < ncc[2]) {
            constexpr int NY = D == 3 ? N * N : N;
            T y[NY];
            mf_cell_part<T, D, K>(P, xs + ((e0 * K) * NB1 + e1 * K) * NB2 + (D == 3 ? e2 * K : 0), NB1 * NB2, NB2, s,
                                  y);
            const T a = cs[ce];
#pragma unroll
            for (int l = 0; l < NY; l++) cr[ce * NL + (D == 3 ? l + N * N * s : l + N * s)] = a * y[l];
        }
    }
    __syncthreads();
    // 3. owned nodes: the results of their cells, ascending, then the epilogue
#pragma unroll
    for (int u = 0; u < NOWN; u++) {
        if (!own[u]) continue;
        if (!PRE_EARLY) pre[u] = epi_load(epi, row[u]);
        const int i = tid + u * BS;
        const int o2 = i % OB2, o1 = (i / OB2) % OB1, o0 = i / (OB1 * OB2);
        if (inner) {
            // the candidate cells of node-box position o + 1, ascending, as in the general path
            const uint4 cw = CAND_EARLY ? candw[u]
                                        : *reinterpret_cast<const uint4*>(
                                              P.cand + 8 * (((o0 + 1) * NB1 + o1 + 1) * NB2 + (D == 3 ? o2 + 1 : 0)));
            const unsigned wv[4] = {cw.x, cw.y, cw.z, cw.w};
            T v[1 << D];
#pragma unroll
            for (int c = 0; c < (1 << D); c++) {
                const unsigned off = (wv[c >> 1] >> (16 * (c & 1))) & 0xffffu;
                const T w = cr[off != 0xffffu ? off : 0];
                v[c] = off != 0xffffu ? w : (T)0;
            }
            T acc = 0;
#pragma unroll
            for (int c = 0; c < (1 << D); c++) acc += v[c];
            if constexpr (EpiTileX<Epi>::value)   // inner tile: no boundary nodes, box position o + 1
                pre[u].x = xs[((o0 + 1) * NB1 + o1 + 1) * NB2 + (D == 3 ? o2 + 1 : 0)];
            epi(row[u], acc, pre[u]);
            continue;
        }
        const int p[3] = {olo[0] + o0, olo[1] + o1, olo[2] + o2};
        bool bnd = false;
#pragma unroll
        for (int a = 0; a < D; a++) bnd |= p[a] == 0 || p[a] == K * n[a];
        T ax;
        if (bnd) {
            ax = x[row[u]];   // identity row of the eliminated system
            if constexpr (EpiTileX<Epi>::value) pre[u].x = ax;   // the tile holds 0 for boundary nodes
        } else {
            // the cells containing p; all lie in the computed box
            int q[3], elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
#pragma unroll
            for (int a = 0; a < 3; a++) {
                q[a] = p[a] - org[a];
                if (a < D) {
                    ehi[a] = q[a] / K;
                    elo[a] = q[a] % K == 0 ? ehi[a] - 1 : ehi[a];
                }
            }
            // 2^d candidates (e = elo + bit), taken in ascending order where e <= ehi; the loop
            // has a fixed trip count so that the LDS reads issue together
            T v[1 << D];
#pragma unroll
            for (int c = 0; c < (1 << D); c++) {
                const int e0 = elo[0] + (c >> (D - 1) & 1), e1 = elo[1] + (c >> (D - 2) & 1);
                const int e2 = D == 3 ? elo[2] + (c & 1) : 0;
                const bool in = e0 <= ehi[0] && e1 <= ehi[1] && (D == 2 || e2 <= ehi[2]);
                const int l = (q[0] - e0 * K) + N * (q[1] - e1 * K) + (D == 3 ? N * N * (q[2] - e2 * K) : 0);
                const int at = in ? ((e0 * CC1 + e1) * CC2 + e2) * NL + l : 0;   // stays inside cr
                const T w = cr[at];
                v[c] = in ? w : (T)0;
            }
            T acc = 0;
#pragma unroll
            for (int c = 0; c < (1 << D); c++) acc += v[c];
            ax = acc;
            if constexpr (EpiTileX<Epi>::value) pre[u].x = xs[(q[0] * NB1 + q[1]) * NB2 + (D == 3 ? q[2] : 0)];
        }
        epi(row[u], ax, pre[u]);
    }
}

// row epilogue launch over a level's tiles; returns false for (d, k) without a kernel
template <typename T, typename Epi>
inline bool launch_mf(hipStream_t st, const MfParams& P, const T* x, const T* coef, const Epi& epi)
{
    long tiles = 1;
    for (int a = 0; a < P.g.d; a++) tiles *= P.ntile[a];
    if (tiles == 0) return true;
    JFEM_CHECK(tiles < (1l << 31), JFEM_ERR_UNSUPPORTED, "too many tiles");
    switch (P.g.d * 10 + P.g.k) {
    case 32: k_mf<T, 3, 2, Epi><<<(unsigned)tiles, MfShape<3, 2>::BS, 0, st>>>(P, x, coef, epi); break;
    case 31: k_mf<T, 3, 1, Epi><<<(unsigned)tiles, MfShape<3, 1>::BS, 0, st>>>(P, x, coef, epi); break;
    case 22: k_mf<T, 2, 2, Epi><<<(unsigned)tiles, MfShape<2, 2>::BS, 0, st>>>(P, x, coef, epi); break;
    case 21: k_mf<T, 2, 1, Epi><<<(unsigned)tiles, MfShape<2, 1>::BS, 0, st>>>(P, x, coef, epi); break;
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

// the inner-tile table cand of k_mf for grid g (see MfParams), 8 entries per node-box position;
// empty when the grid has no inner tile
template <int D, int K>
inline void mf_tables_dk(const Grid& g, std::vector<unsigned short>& cand)
{
    using S = MfShape<D, K>;
    constexpr int N = K + 1, NL = D == 3 ? N * N * N : N * N;
    const int TC[3] = {S::T0, S::T1, D == 3 ? S::T2 : 1};
    const int CC[3] = {S::T0 + 1, S::T1 + 1, D == 3 ? S::T2 + 1 : 1};
    const int NB[3] = {CC[0] * K + 1, CC[1] * K + 1, D == 3 ? CC[2] * K + 1 : 1};
    cand.clear();
    for (int a = 0; a < D; a++)
        if (2 + TC[a] + 1 >= g.n[a]) return;   // no tile with clo >= 2 and clo + T + 1 < n
    const int nnb = NB[0] * NB[1] * NB[2];
    cand.assign((size_t)8 * nnb, 0xffff);
    for (int q0 = 0; q0 < NB[0]; q0++)
        for (int q1 = 0; q1 < NB[1]; q1++)
            for (int q2 = 0; q2 < NB[2]; q2++) {
                const int i = (q0 * NB[1] + q1) * NB[2] + q2;
                // candidate cells of node q: per axis lo = q/K - [q mod K == 0], hi = q/K, the 2^d
                // bit combinations in ascending order (as in k_mf's general path)
                const int q[3] = {q0, q1, q2};
                int elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
                for (int a = 0; a < D; a++) {
                    ehi[a] = q[a] / K;
                    elo[a] = q[a] % K == 0 ? ehi[a] - 1 : ehi[a];
                }
                for (int c = 0; c < (1 << D); c++) {
                    const int e0 = elo[0] + (c >> (D - 1) & 1), e1 = elo[1] + (c >> (D - 2) & 1);
                    const int e2 = D == 3 ? elo[2] + (c & 1) : 0;
                    const bool in = e0 >= 0 && e1 >= 0 && e2 >= 0 && e0 <= ehi[0] && e1 <= ehi[1] &&
                                    (D == 2 || e2 <= ehi[2]) && e0 < CC[0] && e1 < CC[1] && e2 < CC[2];
                    if (!in) continue;
                    const int l = (q[0] - e0 * K) + N * (q[1] - e1 * K) + (D == 3 ? N * N * (q[2] - e2 * K) : 0);
                    cand[(size_t)8 * i + c] = (unsigned short)(((e0 * CC[1] + e1) * CC[2] + e2) * NL + l);
                }
            }
}

inline void mf_tables(const Grid& g, std::vector<unsigned short>& cand)
{
    switch (g.d * 10 + g.k) {
    case 32: mf_tables_dk<3, 2>(g, cand); break;
    case 31: mf_tables_dk<3, 1>(g, cand); break;
    case 22: mf_tables_dk<2, 2>(g, cand); break;
    case 21: mf_tables_dk<2, 1>(g, cand); break;
    default: cand.clear();
    }
}

}  // namespace jfem
