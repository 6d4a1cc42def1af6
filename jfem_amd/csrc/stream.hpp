// CSR-stream sparse row kernel with fused per-row epilogues.
//
// The rows are cut into blocks of consecutive rows whose entries, counted from
// the 16-byte-aligned entry at or below the block's first one, number at most
// ST_CAP (a greedy split of the row pointer, built once per matrix; FP32 values
// use blocks of twice the entries).  A 512-thread workgroup reads its block
// descriptor {first row, end row, first entry, end entry} with one load, then
// issues the row bounds and the epilogue operands of its row and the block's
// values and column indices as 16-byte loads (two rounds of four entries per
// thread in flight), keeps the products val*x[col] in LDS, reduces each row
// with a group of G lanes (G = largest power of two with G * rows <= 512) and
// hands the row sum to an epilogue: plain store (SpMV), residual, Jacobi or
// Chebyshev update, with an optional per-block partial of a dot product.
// Summation order depends only on the matrix, so results are deterministic run
// to run.
#pragma once

#include "common.hpp"

namespace jfem {

#ifndef JFEM_ST_BS
#define JFEM_ST_BS 512
#endif
#ifndef JFEM_ST_CAP
#define JFEM_ST_CAP 4096
#endif
#ifndef JFEM_ST_VEC64
#define JFEM_ST_VEC64 4
#endif
constexpr int ST_BS = JFEM_ST_BS;     // threads per workgroup
constexpr int ST_CAP = JFEM_ST_CAP;   // entries per row block (FP64 values)
#ifndef JFEM_ST_CAP32
#define JFEM_ST_CAP32 (2 * JFEM_ST_CAP)
#endif
constexpr int ST_CAP32 = JFEM_ST_CAP32;   // entries per row block (FP32 values): same bytes per block
static_assert(ST_CAP % ST_BS == 0, "block capacity must be a multiple of the workgroup size");
// consecutive row blocks per XCD (see k_stream).  On the 3D Q2 96^3 matrix, runs of 8-16
// blocks cut the fetched bytes per FP64 product from 6.07 to 5.92 GB and its time from 1.025
// to 0.993 ms; one contiguous range per XCD fetches 5.66 GB but takes 1.065 ms.
constexpr int ST_XCD_RUN = 16;

// entries per staging load: four (one 16-byte column load; values in one or two 16-byte loads)
template <typename VT>
constexpr int st_vec() { return sizeof(VT) == 8 ? JFEM_ST_VEC64 : 4; }

template <int NT>
__device__ inline double st_block_sum(double v, double* sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NT / 64; i++) r += sh[i];
    return r;   // thread 0
}

// vector types of the staging loads
template <typename S, int N>
struct StVec {
    typedef S type __attribute__((ext_vector_type(N)));
};

// Epi must provide:
//   static constexpr bool reduce;
//   struct Pre;  __device__ Pre load(long row) const;   (operands of the row, loaded early)
//   __device__ double operator()(long row, T ax, const Pre&) const
// (the returned value is summed over the block when reduce is true)
template <int CAP, typename VT, typename T, typename Epi>
__global__ __launch_bounds__(ST_BS) void k_stream(const StBlk* __restrict__ blk, const int* __restrict__ rowptr,
                                                  const int* __restrict__ cols, const VT* __restrict__ vals,
                                                  const T* __restrict__ x, Epi epi, double* __restrict__ partials)
{
    constexpr int VEC = st_vec<VT>();              // entries per staging load
    constexpr int R = CAP / (ST_BS * VEC);         // staging loads per thread
    static_assert(R * ST_BS * VEC == CAP, "block capacity must be a multiple of the loads per round");
    typedef typename StVec<VT, VEC>::type vload_t;
    typedef typename StVec<int, VEC>::type cload_t;
    __shared__ T prod[CAP];
    __shared__ double sh[ST_BS / 64];
    const int tid = threadIdx.x;
    // XCD-aware block order.  Workgroups are dealt round-robin to the eight XCDs (each with
    // its own L2), so consecutive workgroups land on different L2s.  Blocks about a dozen
    // apart gather the same x entries (rows of neighbouring node lines), so each XCD gets
    // runs of ST_XCD_RUN consecutive blocks, while all XCDs stay in the same region of the
    // matrix (super-runs of 8 * ST_XCD_RUN blocks).  A bijection on [0, gridDim.x): the tail
    // that does not fill a super-run keeps its order.
    int bid = blockIdx.x;
    {
        constexpr int super = 8 * ST_XCD_RUN;
        const int full = (int)(gridDim.x / super) * super;
        if (bid < full) {
            const int m = bid >> 3;
            bid = (m / ST_XCD_RUN) * super + (bid & 7) * ST_XCD_RUN + (m % ST_XCD_RUN);
        }
    }
    const StBlk B = blk[bid];
    const int nr = B.r1 - B.r0;
    const int a0 = B.p0 & ~(VEC - 1);              // aligned start of the staging window
    double dsum = 0;
    if (B.p1 - a0 > CAP) {
        // one row longer than the block capacity: strided sum over the block
        double acc = 0;
        for (int p = B.p0 + tid; p < B.p1; p += ST_BS) acc += (double)((T)vals[p] * x[cols[p]]);
        const double s = st_block_sum<ST_BS>(acc, sh);
        if (tid == 0) dsum = epi(B.r0, (T)s, epi.load(B.r0));
    } else {
        // the row of this thread's lane group: G lanes per row, G * rows <= ST_BS
        int G = 64;
        while (G > 1 && G * nr > ST_BS) G >>= 1;
        const int j = tid / G, lane = tid & (G - 1);
        int rs = 0, re = 0;
        typename Epi::Pre pre{};
        if (j < nr) {
            rs = rowptr[B.r0 + j] - B.p0;
            re = rowptr[B.r0 + j + 1] - B.p0;
            if (lane == 0) pre = epi.load(B.r0 + j);
        }
        // all staging loads of a thread are issued before any product is formed
        vload_t v[R];
        cload_t c[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int q = a0 + (tid + k * ST_BS) * VEC;
            if (q < B.p1) {
                v[k] = *reinterpret_cast<const vload_t*>(vals + q);
                c[k] = *reinterpret_cast<const cload_t*>(cols + q);
            }
        }
        T xv[R][VEC];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int q = a0 + (tid + k * ST_BS) * VEC;
#pragma unroll
            for (int u = 0; u < VEC; u++)
                if (q + u >= B.p0 && q + u < B.p1) xv[k][u] = x[c[k][u]];
        }
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int q = a0 + (tid + k * ST_BS) * VEC;
#pragma unroll
            for (int u = 0; u < VEC; u++)
                if (q + u >= B.p0 && q + u < B.p1) prod[q + u - B.p0] = (T)v[k][u] * xv[k][u];
        }
        __syncthreads();
        if (j < nr) {
            T acc = 0;
            for (int i = rs + lane; i < re; i += G) acc += prod[i];
            for (int o = G >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
            if (lane == 0) dsum = epi(B.r0 + j, acc, pre);
        }
    }
    if (Epi::reduce) {
        __syncthreads();
        const double s = st_block_sum<ST_BS>(dsum, sh);
        if (tid == 0) partials[bid] = s;
    }
}

template <typename T>
struct EpiStore {
    static constexpr bool reduce = false;
    T* y;
    struct Pre {};
    __device__ Pre load(long) const { return {}; }
    __device__ double operator()(long row, T ax, const Pre&) const { y[row] = ax; return 0.0; }
};

// y = A x and partial of dot(x, y)
struct EpiStoreDot {
    static constexpr bool reduce = true;
    double* y;
    const double* x;
    struct Pre { double x; };
    __device__ Pre load(long row) const { return {x[row]}; }
    __device__ double operator()(long row, double ax, const Pre& p) const { y[row] = ax; return p.x * ax; }
};

template <typename T>
struct EpiResidual {
    static constexpr bool reduce = false;
    const T* b;
    T* r;
    struct Pre { T b; };
    __device__ Pre load(long row) const { return {b[row]}; }
    __device__ double operator()(long row, T ax, const Pre& p) const { r[row] = p.b - ax; return 0.0; }
};

// Jacobi: xn = x + (dinv * (b - A x)) * omega   (JacobiSmoother.java:45-48)
template <typename T>
struct EpiJacobi {
    static constexpr bool reduce = false;
    const T *b, *x, *dinv;
    T* xn;
    T omega;
    struct Pre { T b, x, dinv; };
    __device__ Pre load(long row) const { return {b[row], x[row], dinv[row]}; }
    // x is the product's operand: the matrix-free kernel fills pre.x from its node tile
    static constexpr bool tile_x = true;
    __device__ Pre load_nox(long row) const { return {b[row], (T)0, dinv[row]}; }
    __device__ double operator()(long row, T ax, const Pre& p) const
    {
#pragma clang fp contract(off)
        xn[row] = p.x + (p.dinv * (p.b - ax)) * omega;
        return 0.0;
    }
};

// Chebyshev step: z = dinv*(b - A x); d = first ? z/theta : c1*d + c2*z; xn = x + d
template <typename T>
struct EpiCheby {
    static constexpr bool reduce = false;
    const T *b, *x, *dinv;
    T *xn, *d;
    T theta, c1, c2;
    int first;
    struct Pre { T b, x, dinv, d; };
    __device__ Pre load(long row) const { return {b[row], x[row], dinv[row], first ? (T)0 : d[row]}; }
    static constexpr bool tile_x = true;   // as in EpiJacobi
    __device__ Pre load_nox(long row) const { return {b[row], (T)0, dinv[row], first ? (T)0 : d[row]}; }
    __device__ double operator()(long row, T ax, const Pre& p) const
    {
#pragma clang fp contract(off)
        const T z = p.dinv * (p.b - ax);
        const T dn = first ? z / theta : c1 * p.d + c2 * z;
        d[row] = dn;
        xn[row] = p.x + dn;
        return 0.0;
    }
};

// launch over the matrix's row blocks (FP32 values use blocks of twice the entries)
template <typename VT, typename T, typename Epi>
inline void launch_stream(const jfem_mat* A, const VT* vals, const T* x, const Epi& epi, double* partials)
{
    if constexpr (sizeof(VT) == 4) {
        JFEM_CHECK(A->blk32, JFEM_ERR_STATE, "matrix has no FP32 stream blocks");
        if (A->nrblk32 == 0) return;
        k_stream<ST_CAP32, VT, T, Epi><<<(unsigned)A->nrblk32, ST_BS, 0, A->ctx->stream>>>(A->blk32, A->rowptr, A->colidx,
                                                                                         vals, x, epi, partials);
    } else {
        if (A->nrblk == 0) return;
        k_stream<ST_CAP, VT, T, Epi><<<(unsigned)A->nrblk, ST_BS, 0, A->ctx->stream>>>(A->blk, A->rowptr, A->colidx,
                                                                                     vals, x, epi, partials);
    }
    JFEM_HIP(hipGetLastError());
}

}  // namespace jfem
