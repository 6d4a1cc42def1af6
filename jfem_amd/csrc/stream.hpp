// CSR-stream sparse row kernel with fused per-row epilogues.
//
// The rows are cut into blocks of consecutive rows holding at most ST_CAP
// entries (a greedy split of the row pointer, built once per matrix).  A
// 256-thread workgroup streams its block's values and column indices with
// fully coalesced loads (eight independent loads per thread in flight), keeps
// the products val*x[col] in LDS, then reduces each row with a group of G lanes
// (G = largest power of two with G * rows <= 256) and hands the row sum to an
// epilogue: plain store (SpMV), residual, Jacobi or Chebyshev update, with an
// optional per-block partial of a dot product.  Summation order depends only
// on the matrix, so results are deterministic run to run.
#pragma once

#include "common.hpp"

namespace jfem {

#ifndef JFEM_ST_BS
#define JFEM_ST_BS 256
#endif
#ifndef JFEM_ST_CAP
#define JFEM_ST_CAP 2048
#endif
#ifndef JFEM_ST_NT
#define JFEM_ST_NT 0
#endif
constexpr int ST_BS = JFEM_ST_BS;     // threads per workgroup
constexpr int ST_CAP = JFEM_ST_CAP;   // entries per row block (FP64 values)
constexpr int ST_CAP32 = 2 * ST_CAP;  // entries per row block (FP32 values): same bytes per block
static_assert(ST_CAP % ST_BS == 0, "block capacity must be a multiple of the workgroup size");

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

// Epi must provide:  static constexpr bool reduce;  __device__ double operator()(long row, T ax)
// (the returned value is summed over the block when reduce is true)
template <int CAP, typename VT, typename T, typename Epi>
__global__ __launch_bounds__(ST_BS) void k_stream(const int* __restrict__ rblk, const int* __restrict__ rowptr,
                                                  const int* __restrict__ cols, const VT* __restrict__ vals,
                                                  const T* __restrict__ x, Epi epi, double* __restrict__ partials)
{
    constexpr int ST_PER = CAP / ST_BS;
    __shared__ T prod[CAP];
    __shared__ double sh[ST_BS / 64];
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    const int r0 = rblk[b], r1 = rblk[b + 1];
    const int p0 = rowptr[r0], p1 = rowptr[r1];
    const int nr = r1 - r0, cnt = p1 - p0;
    double dsum = 0;
    if (cnt > CAP) {
        // one row longer than the block capacity: strided sum over the block
        double acc = 0;
        for (int p = p0 + tid; p < p1; p += ST_BS) acc += (double)((T)vals[p] * x[cols[p]]);
        const double s = st_block_sum<ST_BS>(acc, sh);
        if (tid == 0) dsum = epi(r0, (T)s);
    } else {
        // all loads of a thread are issued before any product is formed
        VT v[ST_PER];
        int c[ST_PER];
#pragma unroll
        for (int k = 0; k < ST_PER; k++) {
            const int i = tid + k * ST_BS;
#if JFEM_ST_NT
            if (i < cnt) { v[k] = __builtin_nontemporal_load(&vals[p0 + i]); c[k] = __builtin_nontemporal_load(&cols[p0 + i]); }
#else
            if (i < cnt) { v[k] = vals[p0 + i]; c[k] = cols[p0 + i]; }
#endif
        }
        T xv[ST_PER];
#pragma unroll
        for (int k = 0; k < ST_PER; k++) {
            const int i = tid + k * ST_BS;
            if (i < cnt) xv[k] = x[c[k]];
        }
#pragma unroll
        for (int k = 0; k < ST_PER; k++) {
            const int i = tid + k * ST_BS;
            if (i < cnt) prod[i] = (T)v[k] * xv[k];
        }
        __syncthreads();
        int G = 64;
        while (G > 1 && G * nr > ST_BS) G >>= 1;
        const int lane = tid & (G - 1);
        for (int j = tid / G; j < nr; j += ST_BS / G) {
            const int s = rowptr[r0 + j] - p0, e = rowptr[r0 + j + 1] - p0;
            T acc = 0;
            for (int i = s + lane; i < e; i += G) acc += prod[i];
            for (int o = G >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
            if (lane == 0) dsum += epi(r0 + j, acc);
        }
    }
    if (Epi::reduce) {
        __syncthreads();
        const double s = st_block_sum<ST_BS>(dsum, sh);
        if (tid == 0) partials[b] = s;
    }
}

template <typename T>
struct EpiStore {
    static constexpr bool reduce = false;
    T* y;
    __device__ double operator()(long row, T ax) const { y[row] = ax; return 0.0; }
};

// y = A x and partial of dot(x, y)
struct EpiStoreDot {
    static constexpr bool reduce = true;
    double* y;
    const double* x;
    __device__ double operator()(long row, double ax) const { y[row] = ax; return x[row] * ax; }
};

template <typename T>
struct EpiResidual {
    static constexpr bool reduce = false;
    const T* b;
    T* r;
    __device__ double operator()(long row, T ax) const { r[row] = b[row] - ax; return 0.0; }
};

// Jacobi: xn = x + (dinv * (b - A x)) * omega   (JacobiSmoother.java:45-48)
template <typename T>
struct EpiJacobi {
    static constexpr bool reduce = false;
    const T *b, *x, *dinv;
    T* xn;
    T omega;
    __device__ double operator()(long row, T ax) const
    {
#pragma clang fp contract(off)
        xn[row] = x[row] + (dinv[row] * (b[row] - ax)) * omega;
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
    __device__ double operator()(long row, T ax) const
    {
#pragma clang fp contract(off)
        const T z = dinv[row] * (b[row] - ax);
        const T dn = first ? z / theta : c1 * d[row] + c2 * z;
        d[row] = dn;
        xn[row] = x[row] + dn;
        return 0.0;
    }
};

// launch over the matrix's row blocks (FP32 values use blocks of twice the entries)
template <typename VT, typename T, typename Epi>
inline void launch_stream(const jfem_mat* A, const VT* vals, const T* x, const Epi& epi, double* partials)
{
    if (sizeof(VT) == 4) {
        JFEM_CHECK(A->rblk32, JFEM_ERR_STATE, "matrix has no FP32 stream blocks");
        if (A->nrblk32 == 0) return;
        k_stream<ST_CAP32, VT, T, Epi><<<(unsigned)A->nrblk32, ST_BS, 0, A->ctx->stream>>>(A->rblk32, A->rowptr, A->colidx,
                                                                                         vals, x, epi, partials);
    } else {
        if (A->nrblk == 0) return;
        k_stream<ST_CAP, VT, T, Epi><<<(unsigned)A->nrblk, ST_BS, 0, A->ctx->stream>>>(A->rblk, A->rowptr, A->colidx,
                                                                                     vals, x, epi, partials);
    }
    JFEM_HIP(hipGetLastError());
}

}  // namespace jfem
