// Krylov solvers of IterativeSolver (M/linalg/IterativeSolver.java) and GMRES
// (M/linalg/GMRES.java) on device-resident vectors.  The host keeps the
// scalar control flow; alpha/beta live on the device and are consumed by the
// fused update kernels, so each CG iteration costs one scalar read-back (the
// convergence test ||r|| > tol).  Vector arithmetic keeps the reference's
// operation order per element (x + p*alpha, r - q*alpha, r + p*beta, ...).
#include "common.hpp"

#include <chrono>
#include <cmath>
#include <cstring>

namespace jfem {

constexpr int KB = 256;

template <int NT>
__device__ inline void block_partial(double v, double* sh, double* out)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int i = 0; i < NT / 64; i++) r += sh[i];
        *out = r;
    }
}

// x += alpha p ; r -= alpha q ; partial ||r||^2.  alpha = s[num] / s[den].
__global__ __launch_bounds__(KB) void k_cg_update(long n, const double* __restrict__ sc, int num, int den,
                                                  const double* __restrict__ p, const double* __restrict__ q,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ partials)
{
#pragma clang fp contract(off)
    __shared__ double sh[KB / 64];
    const double alpha = sc[num] / sc[den];
    double acc = 0;
    for (long i = blockIdx.x * (long)KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) {
        x[i] = x[i] + p[i] * alpha;
        const double rn = r[i] - q[i] * alpha;
        r[i] = rn;
        acc += rn * rn;
    }
    block_partial<KB>(acc, sh, &partials[blockIdx.x]);
}

// p = z + p * beta, beta = s[num] / s[den]   (newResiduum.add(defect.mul(beta)))
__global__ __launch_bounds__(KB) void k_xpby(long n, const double* __restrict__ sc, int num, int den,
                                             const double* __restrict__ z, double* __restrict__ p)
{
#pragma clang fp contract(off)
    const double beta = sc[num] / sc[den];
    for (long i = blockIdx.x * (long)KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) p[i] = z[i] + p[i] * beta;
}

// p = z + p * beta with beta = -s[num]/s[den]  (flexible CG, Notay's FCG(1))
__global__ __launch_bounds__(KB) void k_xpby_neg(long n, const double* __restrict__ sc, int num, int den,
                                                 const double* __restrict__ z, double* __restrict__ p)
{
    const double beta = -sc[num] / sc[den];
    for (long i = blockIdx.x * (long)KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) p[i] = z[i] + p[i] * beta;
}

// r = b - t
__global__ void k_sub(long n, const double* __restrict__ b, const double* __restrict__ t, double* __restrict__ r)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) r[i] = b[i] - t[i];
}

// up to RED_SLOTS dots into partials (slot-major), one pass
struct Dots {
    const double* x[RED_SLOTS];
    const double* y[RED_SLOTS];
};
__global__ __launch_bounds__(KB) void k_dots_dev(long n, int m, Dots D, double* __restrict__ partials)
{
    __shared__ double sh[KB / 64];
    for (int s = 0; s < m; s++) {
        double acc = 0;
        for (long i = blockIdx.x * (long)KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) acc += D.x[s][i] * D.y[s][i];
        block_partial<KB>(acc, sh, &partials[s * RED_BLOCKS + blockIdx.x]);
        __syncthreads();
    }
}

static int dots_dev(jfem_ctx* c, long n, int m, const double* const* xs, const double* const* ys, double* dst_dev)
{
    Dots D;
    for (int s = 0; s < m; s++) { D.x[s] = xs[s]; D.y[s] = ys[s]; }
    const int blocks = grid_for(n, KB, RED_BLOCKS);
    k_dots_dev<<<blocks, KB, 0, c->stream>>>(n, m, D, c->red_partials);
    JFEM_HIP(hipGetLastError());
    reduce_partials_dev(c, m, blocks, dst_dev);
    return blocks;
}

static double read_scalar(jfem_ctx* c, const double* dev)
{
    JFEM_HIP(hipMemcpyAsync(c->red_host, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    JFEM_HIP(hipStreamSynchronize(c->stream));
    return c->red_host[0];
}

static void read_scalars(jfem_ctx* c, const double* dev, int m, double* out)
{
    JFEM_HIP(hipMemcpyAsync(c->red_host, dev, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    JFEM_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < m; i++) out[i] = c->red_host[i];
}

static int kgrid(long n) { return grid_for(n, KB, RED_BLOCKS); }

// A x for the solvers: a matrix tagged with a hierarchy's matrix-free finest operator (the
// assembled level matrix, the same operator to rounding) is applied through it, without
// reading its entries; any other matrix through the CSR product
static void op_mul(jfem_mat* A, const double* x, double* y)
{
    if (!mg_op_apply(A, x, y)) spmv(A, x, y);
}

// y = A x and dst = x.y (device)
static void op_mul_dot(jfem_mat* A, const double* x, double* y, double* dst)
{
    if (!mg_op_apply(A, x, y)) {
        spmv_dot_dev(A, x, y, dst);
        return;
    }
    const double* xs[1] = {x};
    const double* ys[1] = {y};
    dots_dev(A->ctx, A->rows, 1, xs, ys, dst);
}

static double true_resnorm(jfem_mat* A, const double* b, const double* x, double* tmp)
{
    jfem_ctx* c = A->ctx;
    op_mul(A, x, tmp);
    k_sub<<<kgrid(A->rows), KB, 0, c->stream>>>(A->rows, b, tmp, tmp);
    return std::sqrt(vec_dot(c, A->rows, tmp, tmp));
}

// ------------------------------------------------------------------ CG / PCG
// scalar slots in ctx->scal: 0 = r.r (or r.z), 1 = p.Ap, 2 = new r.r, 3 = new r.z
static void solve_cg(jfem_mat* A, jfem_mg* M, const double* b, double* x, double tol, long maxit, jfem_solve_info* info,
                     double* hist)
{
    jfem_ctx* c = A->ctx;
    const long n = A->rows;
    double* sc = c->scal;
    double* W = ctx_work(c, 4 * (size_t)n);
    const DevPtr<double> r{W}, z{W + n}, p{W + 2 * n}, q{W + 3 * n};
    // r = b - A x
    op_mul(A, x, q.p);
    k_sub<<<kgrid(n), KB, 0, c->stream>>>(n, b, q.p, r.p);
    if (M) mg_apply(M, r.p, z.p);
    vec_copy(c, p.p, M ? z.p : r.p, n);
    {
        const double* xs[2] = {r.p, r.p};
        const double* ys[2] = {M ? z.p : r.p, r.p};
        dots_dev(c, n, 2, xs, ys, sc);   // sc[0] = r.z, sc[1] = r.r
    }
    double hv[2];
    read_scalars(c, sc, 2, hv);
    double rr = hv[1];
    // the reference's start: sc[0] holds r.z (PCG) or r.r (CG)
    int it = 0;
    const long cap = std::min(n, maxit);
    for (long iter = 0; iter < cap && std::sqrt(rr) > tol; iter++) {
        op_mul_dot(A, p.p, q.p, sc + 1);                   // q = A p, sc[1] = p.Ap
        k_cg_update<<<kgrid(n), KB, 0, c->stream>>>(n, sc, 0, 1, p.p, q.p, x, r.p, c->red_partials);
        reduce_partials_dev(c, 1, kgrid(n), sc + 2);         // sc[2] = r.r (new)
        if (M) {
            mg_apply(M, r.p, z.p);
            const double* xs[1] = {r.p};
            const double* ys[1] = {z.p};
            dots_dev(c, n, 1, xs, ys, sc + 3);               // sc[3] = r.z (new)
            k_xpby<<<kgrid(n), KB, 0, c->stream>>>(n, sc, 3, 0, z.p, p.p);
            JFEM_HIP(hipMemcpyAsync(sc, sc + 3, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        } else {
            k_xpby<<<kgrid(n), KB, 0, c->stream>>>(n, sc, 2, 0, r.p, p.p);
            JFEM_HIP(hipMemcpyAsync(sc, sc + 2, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        }
        rr = read_scalar(c, sc + 2);
        if (hist) hist[it] = std::sqrt(rr);
        it++;
        JFEM_CHECK(std::isfinite(rr), JFEM_ERR_STATE, "CG diverged (non-finite residual)");
    }
    info->iterations = it;
    info->resnorm = std::sqrt(rr);
}

// ------------------------------------------------------------ flexible CG
// Notay, "Flexible conjugate gradients", SIAM J. Sci. Comput. 22 (2000), FCG(1):
// beta = -(z_{k+1} . q_k) / (p_k . q_k), q_k = A p_k.  Outer arithmetic in FP64;
// the preconditioner may run in FP32.
static void solve_fcg(jfem_mat* A, jfem_mg* M, const double* b, double* x, double tol, long maxit,
                      jfem_solve_info* info, double* hist)
{
    jfem_ctx* c = A->ctx;
    const long n = A->rows;
    double* sc = c->scal;
    double* W = ctx_work(c, 4 * (size_t)n);
    const DevPtr<double> r{W}, z{W + n}, p{W + 2 * n}, q{W + 3 * n};
    op_mul(A, x, q.p);
    k_sub<<<kgrid(n), KB, 0, c->stream>>>(n, b, q.p, r.p);
    mg_apply(M, r.p, z.p);
    vec_copy(c, p.p, z.p, n);
    {
        const double* xs[2] = {r.p, r.p};
        const double* ys[2] = {z.p, r.p};
        dots_dev(c, n, 2, xs, ys, sc);   // sc[0] = r.z, sc[1] = r.r
    }
    double hv[2];
    read_scalars(c, sc, 2, hv);
    double rr = hv[1];
    int it = 0;
    const long cap = std::min(n, maxit);
    for (long iter = 0; iter < cap && std::sqrt(rr) > tol; iter++) {
        // alpha = (p . r) / (p . Ap); for FCG(1) p.r = z.r of the current step
        {
            const double* xs[1] = {p.p};
            const double* ys[1] = {r.p};
            dots_dev(c, n, 1, xs, ys, sc + 0);
        }
        op_mul_dot(A, p.p, q.p, sc + 1);
        k_cg_update<<<kgrid(n), KB, 0, c->stream>>>(n, sc, 0, 1, p.p, q.p, x, r.p, c->red_partials);
        reduce_partials_dev(c, 1, kgrid(n), sc + 2);
        mg_apply(M, r.p, z.p);
        {
            const double* xs[1] = {z.p};
            const double* ys[1] = {q.p};
            dots_dev(c, n, 1, xs, ys, sc + 3);   // sc[3] = z.q
        }
        k_xpby_neg<<<kgrid(n), KB, 0, c->stream>>>(n, sc, 3, 1, z.p, p.p);
        const double rr_old = rr;
        rr = read_scalar(c, sc + 2);
        mg_observe_reduction(M, rr_old > 0 ? std::sqrt(rr / rr_old) : 0.0);
        if (hist) hist[it] = std::sqrt(rr);
        it++;
        JFEM_CHECK(std::isfinite(rr), JFEM_ERR_STATE, "FCG diverged (non-finite residual)");
    }
    info->iterations = it;
    info->resnorm = std::sqrt(rr);
}

// --------------------------------------------------------------- BiCGStab
// generic element kernels for BiCGStab
// p = (r + p*beta) - v*(omega*beta)
__global__ void k_bicg_p(long n, double beta, double ob, const double* __restrict__ r, const double* __restrict__ v,
                         double* __restrict__ p)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = (r[i] + p[i] * beta) - v[i] * ob;
}
// s = r - v*alpha
__global__ void k_axmy(long n, double alpha, const double* __restrict__ r, const double* __restrict__ v,
                       double* __restrict__ s)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s[i] = r[i] - v[i] * alpha;
}
// x = (x + a*alpha) + b*omega
__global__ void k_x2(long n, double alpha, double omega, const double* __restrict__ a, const double* __restrict__ bb,
                     double* __restrict__ x)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = (x[i] + a[i] * alpha) + bb[i] * omega;
}

static void solve_gmres(jfem_mat* A, jfem_mg* M, const double* b, double* x, double tol, int restart, int max_cycles,
                        jfem_solve_info* info);

static void solve_bicgstab(jfem_mat* A, jfem_mg* M, const double* b, double* x, double tol, long maxit,
                           jfem_solve_info* info)
{
    jfem_ctx* c = A->ctx;
    const long n = A->rows;
    double* W = ctx_work(c, 10 * (size_t)n);
    const DevPtr<double> v{W}, s{W + n}, t{W + 2 * n}, r0{W + 3 * n}, r{W + 4 * n}, p{W + 5 * n}, y{W + 6 * n},
        z{W + 7 * n}, pt{W + 8 * n}, tmp{W + 9 * n};
    vec_fill(c, v.p, 0, n);
    vec_fill(c, p.p, 0, n);
    op_mul(A, x, tmp.p);
    k_sub<<<kgrid(n), KB, 0, c->stream>>>(n, b, tmp.p, r0.p);
    vec_copy(c, r.p, r0.p, n);
    double alpha = 1, omega = 1, rho = 1, rhoLast = 1, beta = 1;
    int it = 0;
    info->fell_back = 0;
    const long cap = std::min(M ? n * 10 : n, maxit);
    const int g = kgrid(n);
    double rnorm = std::sqrt(vec_dot(c, n, r.p, r.p));
    for (long iter = 0; iter < cap && rnorm > tol; iter++) {
        rho = vec_dot(c, n, r.p, r0.p);
        if (std::fabs(rhoLast) < tol) {
            info->fell_back = 1;
            // the reference restarts from its default start iterate: e_0 (GMRES) or 0 (PGMRES)
            vec_fill(c, x, 0.0, n);
            if (!M) {
                static const double one = 1.0;   // outlives the asynchronous copy
                JFEM_HIP(hipMemcpyAsync(x, &one, sizeof(double), hipMemcpyHostToDevice, c->stream));
            }
            solve_gmres(A, M, b, x, tol, 150, 1000, info);
            info->fell_back = 1;
            return;
        }
        beta = alpha / omega * rho / rhoLast;
        k_bicg_p<<<g, KB, 0, c->stream>>>(n, beta, omega * beta, r.p, v.p, p.p);
        if (M) { mg_apply(M, p.p, y.p); op_mul(A, y.p, v.p); } else op_mul(A, p.p, v.p);
        alpha = rho / vec_dot(c, n, v.p, r0.p);
        k_axmy<<<g, KB, 0, c->stream>>>(n, alpha, r.p, v.p, s.p);
        if (std::sqrt(vec_dot(c, n, s.p, s.p)) < tol) {
            k_x2<<<g, KB, 0, c->stream>>>(n, alpha, 0.0, M ? y.p : p.p, M ? y.p : p.p, x);
            break;
        }
        if (M) {
            mg_apply(M, s.p, z.p);
            op_mul(A, z.p, t.p);
            mg_apply(M, t.p, pt.p);
            double d[2];
            const double* xs[2] = {pt.p, pt.p};
            const double* ys[2] = {z.p, pt.p};
            vec_dots(c, n, 2, xs, ys, d);
            omega = d[0] / d[1];
            k_x2<<<g, KB, 0, c->stream>>>(n, alpha, omega, y.p, z.p, x);
        } else {
            op_mul(A, s.p, t.p);
            double d[2];
            const double* xs[2] = {t.p, t.p};
            const double* ys[2] = {s.p, t.p};
            vec_dots(c, n, 2, xs, ys, d);
            omega = d[0] / d[1];
            k_x2<<<g, KB, 0, c->stream>>>(n, alpha, omega, p.p, s.p, x);
        }
        k_axmy<<<g, KB, 0, c->stream>>>(n, omega, s.p, t.p, r.p);
        rhoLast = rho;
        it++;
        rnorm = std::sqrt(vec_dot(c, n, r.p, r.p));
        JFEM_CHECK(std::isfinite(rnorm), JFEM_ERR_STATE, "BiCGStab diverged (non-finite residual)");
    }
    info->iterations = it;
    info->resnorm = true_resnorm(A, b, x, tmp.p);
}

// ------------------------------------------------------------------- GMRES
// Java 11 DoubleStream.sum() (compensated; final sum + compensation)
struct JSum {
    double s = 0, c = 0, simple = 0;
    void add(double v)
    {
#pragma clang fp contract(off)
        const double tmp = v - c;
        const double sum = s;
        const double velvel = sum + tmp;
        c = (velvel - sum) - tmp;
        s = velvel;
        simple += v;
    }
    double final_() const
    {
        const double tmp = s + c;
        if (std::isnan(tmp) && std::isinf(simple)) return simple;
        return tmp;
    }
};

// multi-dot f[j] = V_j . q for j < nv (partials [j][block]); eight vectors per
// sweep so that q is read nv / 8 times (each dot keeps its per-thread order)
__global__ __launch_bounds__(KB) void k_multidot(long n, int nv, const double* __restrict__ V, long ldv,
                                                 const double* __restrict__ q, double* __restrict__ partials, int nb)
{
    constexpr int CH = 8;
    __shared__ double sh[KB / 64];
    for (int j0 = 0; j0 < nv; j0 += CH) {
        const int m = nv - j0 < CH ? nv - j0 : CH;
        double acc[CH];
#pragma unroll
        for (int t = 0; t < CH; t++) acc[t] = 0;
        for (long i = blockIdx.x * (long)KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) {
            const double qi = q[i];
#pragma unroll
            for (int t = 0; t < CH; t++)
                if (t < m) acc[t] += V[(long)(j0 + t) * ldv + i] * qi;
        }
        for (int t = 0; t < m; t++) {
            block_partial<KB>(acc[t], sh, &partials[(long)(j0 + t) * nb + blockIdx.x]);
            __syncthreads();
        }
    }
}

// out[j] = sum of the nb partials of vector j; one workgroup per vector
__global__ __launch_bounds__(KB) void k_sum_partials(int nv, int nb, const double* __restrict__ partials,
                                                     double* __restrict__ out)
{
    __shared__ double sh[KB / 64];
    const int j = blockIdx.x;
    if (j >= nv) return;
    double acc = 0;
    for (int i = threadIdx.x; i < nb; i += KB) acc += partials[(long)j * nb + i];
    block_partial<KB>(acc, sh, &out[j]);
}

// w = q - sum_j V_j f_j   (orthogonalPartInSubspace summed left to right from 0)
__global__ void k_gs_sub(long n, int nv, const double* __restrict__ V, long ldv, const double* __restrict__ f,
                         const double* __restrict__ q, double* __restrict__ w)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double acc = 0;
        for (int j = 0; j < nv; j++) acc = acc + V[(long)j * ldv + i] * f[j];
        w[i] = q[i] - acc;
    }
}

// out = base + sum_j V_j a_j
__global__ void k_lincomb(long n, int nv, const double* __restrict__ V, long ldv, const double* __restrict__ a,
                          double* __restrict__ x)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double acc = 0;
        for (int j = 0; j < nv; j++) acc = acc + V[(long)j * ldv + i] * a[j];
        x[i] = x[i] + acc;
    }
}

__global__ void k_scale(long n, double s, const double* __restrict__ a, double* __restrict__ b)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i] * s;
}

static void solve_gmres(jfem_mat* A, jfem_mg* M, const double* b, double* x, double tol, int restart, int max_cycles,
                        jfem_solve_info* info)
{
#pragma clang fp contract(off)
    jfem_ctx* c = A->ctx;
    const long n = A->rows;
    JFEM_CHECK(restart >= 1 && restart < 100000, JFEM_ERR_ARG, "bad restart length");
    const int ld = restart + 10;
    const int maxv = restart + 3;
    const long ldv = n;
    double* V = nullptr;
    JFEM_CHECK(maxv * (double)n * 8 < 200e9, JFEM_ERR_NOMEM, "Krylov basis exceeds the device memory budget");
    JFEM_HIP(hipMalloc(&V, sizeof(double) * (size_t)maxv * n));
    DevBuf<double> t(n), r0(n), q(n), tq(n), w(n), fdev(maxv), part((long)maxv * RED_BLOCKS);
    std::vector<double> f(maxv), H((size_t)ld * ld), gamma(ld), cs(ld), sn(ld), alpha(ld);
    int restarts = 0, iterations = 0, cycles = 0;
    const int nb = kgrid(n);
    double residual = 0;
    try {
        for (;;) {
            // r = M (b - A x)
            op_mul(A, x, t.p);
            k_sub<<<nb, KB, 0, c->stream>>>(n, b, t.p, r0.p);
            double* r = V;   // v_0 slot holds r, normalised in place
            if (M) mg_apply(M, r0.p, r); else vec_copy(c, r, r0.p, n);
            const double rn = std::sqrt(vec_dot(c, n, r, r));
            if (rn <= tol) { residual = -1; break; }
            std::fill(H.begin(), H.end(), 0.0);
            std::fill(gamma.begin(), gamma.end(), 0.0);
            gamma[0] = rn;
            k_scale<<<nb, KB, 0, c->stream>>>(n, 1. / gamma[0], r, V);
            int nv = 1, j;
            for (j = 0; j < n; j++) {
                const double* vj = V + (long)j * ldv;
                if (M) { op_mul(A, vj, tq.p); mg_apply(M, tq.p, q.p); } else op_mul(A, vj, q.p);
                k_multidot<<<nb, KB, 0, c->stream>>>(n, nv, V, ldv, q.p, part.p, nb);
                k_sum_partials<<<nv, KB, 0, c->stream>>>(nv, nb, part.p, fdev.p);
                k_gs_sub<<<nb, KB, 0, c->stream>>>(n, nv, V, ldv, fdev.p, q.p, w.p);
                JFEM_HIP(hipMemcpyAsync(f.data(), fdev.p, sizeof(double) * nv, hipMemcpyDeviceToHost, c->stream));
                const double wn = std::sqrt(vec_dot(c, n, w.p, w.p));   // synchronises
                for (int i = 0; i < nv; i++) if (f[i] != 0) H[(size_t)i * ld + j] = H[(size_t)i * ld + j] + f[i];
                if (wn != 0) H[(size_t)(j + 1) * ld + j] = H[(size_t)(j + 1) * ld + j] + wn;
                for (int i = 0; i < j; i++) {   // calculateGivens with set() = append(value - at)
                    const double hij = H[(size_t)i * ld + j], hipj = H[(size_t)(i + 1) * ld + j];
                    const double a1 = hij * cs[i] + hipj * sn[i];
                    H[(size_t)i * ld + j] = H[(size_t)i * ld + j] + (a1 - H[(size_t)i * ld + j]);
                    const double a2 = -hij * sn[i] + hipj * cs[i];
                    H[(size_t)(i + 1) * ld + j] = H[(size_t)(i + 1) * ld + j] + (a2 - H[(size_t)(i + 1) * ld + j]);
                }
                const double hjj = H[(size_t)j * ld + j], hj1 = H[(size_t)(j + 1) * ld + j];
                const double beta = std::sqrt(std::pow(hjj, 2) + std::pow(hj1, 2));
                sn[j] = hj1 / beta;
                cs[j] = hjj / beta;
                H[(size_t)j * ld + j] = H[(size_t)j * ld + j] + (beta - H[(size_t)j * ld + j]);
                gamma[j + 1] = -sn[j] * gamma[j];
                gamma[j] = cs[j] * gamma[j];
                if (std::fabs(gamma[j + 1]) < tol || j > restart) break;
                JFEM_CHECK(nv < maxv, JFEM_ERR_STATE, "Krylov basis overflow");
                k_scale<<<nb, KB, 0, c->stream>>>(n, 1. / H[(size_t)(j + 1) * ld + j], w.p, V + (long)nv * ldv);
                nv++;
                iterations++;
            }
            if (j == n) j--;
            for (int i = j; i >= 0; i--) {   // calculateAlpha
                JSum s;
                for (int k2 = i + 1; k2 <= j; k2++) s.add(H[(size_t)i * ld + k2] * alpha[k2]);
                alpha[i] = 1. / H[(size_t)i * ld + i] * (gamma[i] - s.final_());
            }
            JFEM_HIP(hipMemcpyAsync(fdev.p, alpha.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, c->stream));
            k_lincomb<<<nb, KB, 0, c->stream>>>(n, j + 1, V, ldv, fdev.p, x);
            residual = true_resnorm(A, b, x, t.p);
            cycles++;
            const bool again = (j > restart && restarts < 200) || residual > tol;
            if (!again || cycles >= max_cycles) break;
            restarts++;
        }
    } catch (...) {
        (void)hipFree(V);
        throw;
    }
    (void)hipFree(V);
    info->iterations = iterations;
    info->resnorm = residual >= 0 ? residual : true_resnorm(A, b, x, t.p);
}


// ------------------------------------------------------------------- smoothers on any matrix
// 1 / D_ii with |D_ii| < 1e-10 taken as 1 (JacobiSmoother: getDiagonalMatrix, invert)
__global__ void k_dinv_of(long n, const double* __restrict__ d, double* __restrict__ dinv)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = std::fabs(d[i]) < 1e-10 ? 1.0 : d[i];
        dinv[i] = 1.0 / v;
    }
}

// x = x + (dinv r) omega (Jacobi) or x = x + r omega (Richardson, dinv null)
__global__ void k_relax(long n, double omega, const double* __restrict__ dinv, const double* __restrict__ r,
                        double* __restrict__ x)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = x[i] + (dinv ? dinv[i] * r[i] : r[i]) * omega;
}

// Chebyshev step: d = (dinv r) / theta (first) or d = c1 d + c2 (dinv r); x = x + d
__global__ void k_cheb_apply(long n, int first, double theta, double c1, double c2, const double* __restrict__ dinv,
                             const double* __restrict__ r, double* __restrict__ d, double* __restrict__ x)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = first ? (dinv[i] * r[i]) / theta : c1 * d[i] + c2 * (dinv[i] * r[i]);
        d[i] = v;
        x[i] = x[i] + v;
    }
}

// work: 4 n doubles (A x, the residual, D^-1, the Chebyshev direction); every element a sweep
// reads has been written earlier in the same call, so the contents on entry do not matter
static void smooth_any(jfem_mat* A, int kind, int runs, double omega, int degree, double lmin, double lmax,
                       const double* b, double* x, double* work)
{
    jfem_ctx* c = A->ctx;
    const long n = A->rows;
    const int nb = kgrid(n);
    const DevPtr<double> t{work}, r{work + n}, dinv{work + 2 * n}, dd{work + 3 * n};
    if (kind != JFEM_SM_RICHARDSON) {
        mat_diag(A, t);
        k_dinv_of<<<nb, KB, 0, c->stream>>>(n, t, dinv);
    }
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    for (int it = 0; it < runs; it++) {
        if (kind == JFEM_SM_CHEBYSHEV) {
            double rho_old = 1.0 / sigma;
            for (int kd = 0; kd < degree; kd++) {
                op_mul(A, x, t);
                k_sub<<<nb, KB, 0, c->stream>>>(n, b, t, r);
                double c1 = 0, c2 = 0;
                if (kd > 0) {
                    const double rho = 1.0 / (2.0 * sigma - rho_old);
                    c1 = rho * rho_old;
                    c2 = 2.0 * rho / delta;
                    rho_old = rho;
                }
                k_cheb_apply<<<nb, KB, 0, c->stream>>>(n, kd == 0, theta, c1, c2, dinv, r, dd, x);
            }
            continue;
        }
        op_mul(A, x, t);
        k_sub<<<nb, KB, 0, c->stream>>>(n, b, t, r);
        if (kind == JFEM_SM_RICHARDSON && std::sqrt(vec_dot(c, n, r, r)) < 1e-10) break;   // RichardsonSmoother :36-41
        k_relax<<<nb, KB, 0, c->stream>>>(n, omega, kind == JFEM_SM_JACOBI ? dinv.p : nullptr, r, x);
    }
    JFEM_HIP(hipGetLastError());
    JFEM_HIP(hipStreamSynchronize(c->stream));
}

// ------------------------------------------------------------------- GMRES2
// GMRES2.solveInternal (M/linalg/GMRES2.java:59-104) with Arnoldi (M/linalg/Arnoldi.java:13-82):
// minimal-residual steps x += t v along MA-orthonormalised directions (MA = M(A .)) of the
// power sequence q_{k+1} = MA (q_k / |q_k|); the residual and the preconditioned residual
// are updated recursively, the stop test is |r| <= tol (absolute), and a direction whose
// |MA w| falls below 1e-13 after orthogonalisation restarts from the current iterate.
// The same operation sequence as the reference (y + x a forms as axpby with b = 1); the
// dot products are the library's deterministic two-stage sums.  The reference keeps every
// direction; here at most `cap` are held (device memory), reaching it restarts, and maxit
// bounds the iterations.
static void ma_apply(jfem_mat* A, jfem_mg* M, const double* x, double* tmp, double* y)
{
    op_mul(A, x, M ? tmp : y);
    if (M) mg_apply(M, tmp, y);
}

static void solve_gmres2(jfem_mat* A, jfem_mg* M, const double* b, double* x, double tol, long maxit,
                         jfem_solve_info* info)
{
#pragma clang fp contract(off)
    jfem_ctx* c = A->ctx;
    const long n = A->rows;
    long cap = maxit + 2 < 512 ? maxit + 2 : 512;
    while (cap > 8 && 2.0 * cap * n * 8 > 100e9) cap /= 2;
    DevBuf<double> r(n), Mr(n), t(n), q(n), qn(n), w(n), aw(n), V((size_t)cap * n), AV((size_t)cap * n);
    const int nb = kgrid(n);
    long it = 0;
    for (;;) {
        op_mul(A, x, t);
        k_sub<<<nb, KB, 0, c->stream>>>(n, b, t, r);
        if (M) mg_apply(M, r, Mr); else vec_copy(c, Mr, r, n);
        // Arnoldi constructor
        k_scale<<<nb, KB, 0, c->stream>>>(n, 1. / std::sqrt(vec_dot(c, n, Mr, Mr)), Mr, w);
        ma_apply(A, M, w, t, q);
        const double n1 = std::sqrt(vec_dot(c, n, q, q));
        k_scale<<<nb, KB, 0, c->stream>>>(n, 1. / n1, w, V.p);
        k_scale<<<nb, KB, 0, c->stream>>>(n, 1. / n1, q, AV.p);
        long nv = 1;
        k_scale<<<nb, KB, 0, c->stream>>>(n, 1. / n1, q, w);
        ma_apply(A, M, w, t, qn);
        vec_copy(c, q, qn, n);
        bool restart = false;
        while (std::sqrt(vec_dot(c, n, r, r)) > tol) {
            if (it >= maxit) break;
            const double* v = V.p + (size_t)(nv - 1) * n;
            const double* av = AV.p + (size_t)(nv - 1) * n;
            const double tt = vec_dot(c, n, Mr, av) / vec_dot(c, n, av, av);
            vec_axpby(c, n, tt, v, 1.0, x);
            op_mul(A, v, t);
            vec_axpby(c, n, -tt, t, 1.0, r);
            vec_axpby(c, n, -tt, av, 1.0, Mr);
            it++;
            // arnoldiStep
            k_scale<<<nb, KB, 0, c->stream>>>(n, 1. / std::sqrt(vec_dot(c, n, q, q)), q, w);
            ma_apply(A, M, w, t, qn);
            vec_copy(c, aw, qn, n);
            for (long i = 0; i < nv; i++) {
                const double p = vec_dot(c, n, aw, AV.p + (size_t)i * n);
                vec_axpby(c, n, -p, V.p + (size_t)i * n, 1.0, w);
                vec_axpby(c, n, -p, AV.p + (size_t)i * n, 1.0, aw);
            }
            const double na = std::sqrt(vec_dot(c, n, aw, aw));
            vec_copy(c, q, qn, n);
            if (na < 1e-13 || nv == cap) { restart = true; break; }
            k_scale<<<nb, KB, 0, c->stream>>>(n, 1. / na, w, V.p + (size_t)nv * n);
            k_scale<<<nb, KB, 0, c->stream>>>(n, 1. / na, aw, AV.p + (size_t)nv * n);
            nv++;
        }
        JFEM_HIP(hipGetLastError());
        if (!restart || it >= maxit) break;
    }
    info->iterations = (int)it;
    info->resnorm = true_resnorm(A, b, x, t);
}

// ------------------------------------------------- Stokes: block-preconditioned FGMRES
// velocity component c of a mixed vector (node order) <-> scalar space vector (DOF order)
__global__ void k_gather_rank(long n, const int* __restrict__ rank, const double* __restrict__ src,
                              double* __restrict__ dst)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[rank[i]];
}

// dst[rank[i]] = z[i], or r[i] on a Dirichlet row (the identity block of the eliminated system)
__global__ void k_scatter_rank(long n, const int* __restrict__ rank, const double* __restrict__ z,
                               const double* __restrict__ r, const unsigned char* __restrict__ bnd,
                               double* __restrict__ dst)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[rank[i]] = bnd && bnd[i] ? r[i] : z[i];
}

// z_p = s * (B z_u - r_p)
__global__ void k_schur_diag(long n, const double* __restrict__ s, const double* __restrict__ bz,
                             const double* __restrict__ r, double* __restrict__ z)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        z[i] = s[i] * (bz[i] - r[i]);
}

// t = B z_u - r_p (in place in bz)
__global__ void k_sub_inplace(long n, const double* __restrict__ r, double* __restrict__ bz)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        bz[i] = bz[i] - r[i];
}

// lumped-mass inverse in DOF order: pinv[i] = s[rank[i]] for free rows, 0 for pinned ones (s < 0)
__global__ void k_gather_pos(long n, const int* __restrict__ rank, const double* __restrict__ s,
                             double* __restrict__ pinv)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = s[rank[i]];
        pinv[i] = v > 0 ? v : 0.0;
    }
}

// first Chebyshev step from y = 0: d = pinv * t / theta, y = d
__global__ void k_mass_cheb0(long n, double inv_theta, const double* __restrict__ pinv, const double* __restrict__ t,
                             double* __restrict__ d, double* __restrict__ y)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = pinv[i] * t[i] * inv_theta;
        d[i] = v;
        y[i] = v;
    }
}

// d = a d + b pinv (t - q), y += d   (q = M y)
__global__ void k_mass_cheb(long n, double a, double b, const double* __restrict__ pinv, const double* __restrict__ t,
                            const double* __restrict__ q, double* __restrict__ d, double* __restrict__ y)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = a * d[i] + b * pinv[i] * (t[i] - q[i]);
        d[i] = v;
        y[i] += v;
    }
}

// node order <- DOF order, pinned rows (s < 0) keep the identity block: z = -t
__global__ void k_scatter_mass(long n, const int* __restrict__ rank, const double* __restrict__ y,
                               const double* __restrict__ s, const double* __restrict__ t, double* __restrict__ z)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long j = rank[i];
        z[j] = s[j] < 0 ? -t[j] : y[i];
    }
}

__global__ void k_row_slice(long n, const int* __restrict__ rowptr, long first, int* __restrict__ out)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i <= n; i += (long)gridDim.x * blockDim.x)
        out[i] = rowptr[first + i] - rowptr[first];
}

struct StokesPrec {
    jfem_mat* A = nullptr;
    jfem_mg* M = nullptr;
    int d = 0;
    long nv = 0, np = 0;
    const double* ps = nullptr;
    jfem_mat* B = nullptr;       // copy of the pressure rows of A (np rows over all n columns)
    DevBuf<int> rank;            // lexicographic node rank of every velocity DOF
    DevBuf<double> rs[MAXD], zs[MAXD], bz;
    jfem_mg* Mc[MAXD] = {nullptr, nullptr, nullptr};   // hierarchy of each velocity component
    hipEvent_t ev_in = nullptr, ev_done[MAXD] = {nullptr, nullptr, nullptr};
    // optional Schur complement solve: Chebyshev steps on the pressure mass matrix (DOF order)
    jfem_mat* Mp = nullptr;
    int mits = 0;
    double mlo = 0, mhi = 1;
    DevBuf<int> prank;           // lexicographic node rank of every pressure DOF
    DevBuf<double> pinv, tp, yp, dp, qp;
    StokesPrec(jfem_mat* A_, jfem_mg* M_, const double* ps_, long np_, const jfem_stokes_opts* o)
        : A(A_), M(M_), np(np_), ps(ps_)
    {
        try {
            init(o);
        } catch (...) {
            release();
            throw;
        }
    }
    void release()
    {
        mat_free(B);
        B = nullptr;
        if (ev_in) (void)hipEventDestroy(ev_in);
        ev_in = nullptr;
        for (auto& e : ev_done) {
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
        }
    }
    void init(const jfem_stokes_opts* o)
    {
        nv = mg_size(M);
        const long n = A->rows;
        JFEM_CHECK(nv > 0 && np > 0 && (n - np) % nv == 0, JFEM_ERR_ARG, "system size is not dim * velocity + pressure");
        d = (int)((n - np) / nv);
        const Grid& g = mg_finest_grid(M);
        JFEM_CHECK(d == g.d, JFEM_ERR_ARG, "system size is not dim * velocity + pressure");
        jfem_ctx* c = A->ctx;
        rank.alloc(nv);
        bz.alloc(np);
        for (int k = 0; k < d; k++) {
            Mc[k] = M;
            if (k > 0 && o && o->component_mg[k - 1]) {
                jfem_mg* Mk = o->component_mg[k - 1];
                const Grid& gk = mg_finest_grid(Mk);
                bool same = mg_size(Mk) == nv && gk.d == g.d && gk.k == g.k;
                for (int a = 0; a < g.d; a++) same = same && gk.n[a] == g.n[a];
                JFEM_CHECK(same, JFEM_ERR_ARG, "component hierarchy differs from the velocity hierarchy");
                Mc[k] = Mk;
            }
            // a component on another stream gets its own work vectors
            if (k == 0 || mg_stream(Mc[k]) != c->stream) {
                rs[k].alloc(nv);
                zs[k].alloc(nv);
            }
        }
        // components on other streams than the solver's wait for the residual and are waited for
        for (int k = 0; k < d; k++)
            if (mg_stream(Mc[k]) != c->stream) {
                if (!ev_in) JFEM_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
                JFEM_HIP(hipEventCreateWithFlags(&ev_done[k], hipEventDisableTiming));
            }
        jfem_space vs;
        vs.ctx = c;
        vs.grid = g;
        space_lex_rank(&vs, rank);
        const long first = d * nv;
        int base = 0;
        JFEM_HIP(hipMemcpyAsync(&base, A->rowptr + first, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        JFEM_HIP(hipStreamSynchronize(c->stream));
        B = new jfem_mat();
        {
            B->ctx = c;
            B->rows = np;
            B->cols = n;
            B->nnz = A->nnz - base;
            JFEM_HIP(hipMalloc(&B->rowptr, sizeof(int) * (np + 1)));
            JFEM_HIP(hipMalloc(&B->colidx, sizeof(int) * (B->nnz + NNZ_PAD)));
            JFEM_HIP(hipMalloc(&B->vals, sizeof(double) * (B->nnz + NNZ_PAD)));
            k_row_slice<<<grid_for(np + 1, 256, 4096), 256, 0, c->stream>>>(np, A->rowptr, first, B->rowptr);
            JFEM_HIP(hipGetLastError());
            JFEM_HIP(hipMemcpyAsync(B->colidx, A->colidx + base, sizeof(int) * B->nnz, hipMemcpyDeviceToDevice, c->stream));
            JFEM_HIP(hipMemcpyAsync(B->vals, A->vals + base, sizeof(double) * B->nnz, hipMemcpyDeviceToDevice, c->stream));
            B->lanes = choose_lanes((double)B->nnz / np);
            mat_build_blocks(B);
        }
        if (o && o->pressure_mass && o->mass_its > 0) {
            Mp = o->pressure_mass;
            mits = o->mass_its;
            mlo = o->mass_lmin;
            mhi = o->mass_lmax;
            JFEM_CHECK(Mp->rows == np && Mp->cols == np, JFEM_ERR_ARG, "pressure mass matrix has the wrong size");
            JFEM_CHECK(mlo > 0 && mhi > mlo, JFEM_ERR_ARG, "bad mass spectrum bounds");
            prank.alloc(np);
            pinv.alloc(np);
            tp.alloc(np);
            yp.alloc(np);
            dp.alloc(np);
            qp.alloc(np);
            jfem_space ps_space;
            ps_space.ctx = c;
            ps_space.grid = g;
            ps_space.grid.k = g.k - 1;
            JFEM_CHECK(ps_space.grid.k >= 1 && ps_space.grid.ndofs() == np, JFEM_ERR_ARG,
                       "pressure space is not the Q(k-1) space of the velocity grid");
            space_lex_rank(&ps_space, prank);
            k_gather_pos<<<kgrid(np), KB, 0, c->stream>>>(np, prank, ps, pinv);
            JFEM_HIP(hipGetLastError());
        }
    }
    ~StokesPrec() { release(); }
    // z = P^-1 r
    void apply(const double* r, double* z)
    {
        jfem_ctx* c = A->ctx;
        const int gv = kgrid(nv);
        if (ev_in) JFEM_HIP(hipEventRecord(ev_in, c->stream));
        for (int k = 0; k < d; k++) {
            const hipStream_t sk = mg_stream(Mc[k]);
            double* rk = rs[k].p ? rs[k].p : rs[0].p;
            double* zk = zs[k].p ? zs[k].p : zs[0].p;
            if (ev_done[k]) JFEM_HIP(hipStreamWaitEvent(sk, ev_in, 0));
            k_gather_rank<<<gv, KB, 0, sk>>>(nv, rank, r + k * nv, rk);
            mg_apply(Mc[k], rk, zk);
            k_scatter_rank<<<gv, KB, 0, sk>>>(nv, rank, zk, rk, mg_finest_bnd(Mc[k]), z + k * nv);
            if (ev_done[k]) JFEM_HIP(hipEventRecord(ev_done[k], sk));
        }
        for (int k = 0; k < d; k++)
            if (ev_done[k]) JFEM_HIP(hipStreamWaitEvent(c->stream, ev_done[k], 0));
        vec_fill(c, z + d * nv, 0.0, np);
        spmv(B, z, bz);
        if (!Mp) {
            k_schur_diag<<<kgrid(np), KB, 0, c->stream>>>(np, ps, bz, r + d * nv, z + d * nv);
            JFEM_HIP(hipGetLastError());
            return;
        }
        // z_p = M^-1 (B z_u - r_p) by Chebyshev iteration (Saad, Iterative Methods, Alg. 12.1)
        // preconditioned by the lumped mass, on [mlo, mhi]; pinned rows keep z = -t
        const int gp = kgrid(np);
        k_sub_inplace<<<gp, KB, 0, c->stream>>>(np, r + d * nv, bz);
        k_gather_rank<<<gp, KB, 0, c->stream>>>(np, prank, bz, tp);
        const double theta = 0.5 * (mhi + mlo), delta = 0.5 * (mhi - mlo), sigma = theta / delta;
        k_mass_cheb0<<<gp, KB, 0, c->stream>>>(np, 1.0 / theta, pinv, tp, dp, yp);
        double rho = 1.0 / sigma;
        for (int k = 1; k < mits; k++) {
            spmv(Mp, yp, qp);
            const double rn = 1.0 / (2.0 * sigma - rho);
            k_mass_cheb<<<gp, KB, 0, c->stream>>>(np, rn * rho, 2.0 * rn / delta, pinv, tp, qp, dp, yp);
            rho = rn;
        }
        k_scatter_mass<<<gp, KB, 0, c->stream>>>(np, prank, yp, ps, bz, z + d * nv);
        JFEM_HIP(hipGetLastError());
    }
};

// right-preconditioned flexible GMRES (Saad 1993) with classical Gram-Schmidt
// applied twice and Givens rotations
static void solve_stokes_fgmres(jfem_mat* A, StokesPrec& P, const double* b, double* x, double rtol, int restart,
                                long maxit, jfem_solve_info* info, double* hist)
{
#pragma clang fp contract(off)
    jfem_ctx* c = A->ctx;
    const long n = A->rows;
    JFEM_CHECK(restart >= 1 && restart <= 1000, JFEM_ERR_ARG, "bad restart length");
    const int m = restart;
    JFEM_CHECK((2.0 * m + 1) * n * 8 < 200e9, JFEM_ERR_NOMEM, "Krylov basis exceeds the device memory budget");
    DevBuf<double> V((size_t)(m + 1) * n), Z((size_t)m * n), w(n), q(n), t(n), fdev(m + 1),
        part((size_t)(m + 1) * RED_BLOCKS);
    const int nb = kgrid(n);
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), f(m + 1), f2(m + 1);
    auto Hij = [&](int i, int j) -> double& { return H[(size_t)i * m + j]; };
    const double bn = std::sqrt(vec_dot(c, n, b, b));
    long it = 0;
    double res = 0;
    if (bn == 0) {
        vec_fill(c, x, 0.0, n);
        info->iterations = 0;
        info->resnorm = 0;
        return;
    }
    const double target = rtol * bn;
    for (;;) {
        op_mul(A, x, t);
        k_sub<<<nb, KB, 0, c->stream>>>(n, b, t, w);
        JFEM_HIP(hipGetLastError());
        res = std::sqrt(vec_dot(c, n, w, w));
        if (res <= target || it >= maxit) break;
        k_scale<<<nb, KB, 0, c->stream>>>(n, 1.0 / res, w, V.p);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = res;
        int jn = 0;   // columns of this cycle
        bool stop = false;
        for (int j = 0; j < m && !stop; j++) {
            double* vj = V.p + (size_t)j * n;
            double* zj = Z.p + (size_t)j * n;
            P.apply(vj, zj);
            op_mul(A, zj, q);
            // h = V^T q, w = q - V h; h2 = V^T w, q = w - V h2
            k_multidot<<<nb, KB, 0, c->stream>>>(n, j + 1, V.p, n, q, part, nb);
            k_sum_partials<<<j + 1, KB, 0, c->stream>>>(j + 1, nb, part, fdev);
            k_gs_sub<<<nb, KB, 0, c->stream>>>(n, j + 1, V.p, n, fdev, q, w);
            JFEM_HIP(hipMemcpyAsync(f.data(), fdev, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, c->stream));
            k_multidot<<<nb, KB, 0, c->stream>>>(n, j + 1, V.p, n, w, part, nb);
            k_sum_partials<<<j + 1, KB, 0, c->stream>>>(j + 1, nb, part, fdev);
            k_gs_sub<<<nb, KB, 0, c->stream>>>(n, j + 1, V.p, n, fdev, w, q);
            JFEM_HIP(hipMemcpyAsync(f2.data(), fdev, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, c->stream));
            JFEM_HIP(hipGetLastError());
            const double hn = std::sqrt(vec_dot(c, n, q, q));   // synchronises
            for (int i = 0; i <= j; i++) Hij(i, j) = f[i] + f2[i];
            Hij(j + 1, j) = hn;
            for (int i = 0; i < j; i++) {
                const double a1 = cs[i] * Hij(i, j) + sn[i] * Hij(i + 1, j);
                const double a2 = -sn[i] * Hij(i, j) + cs[i] * Hij(i + 1, j);
                Hij(i, j) = a1;
                Hij(i + 1, j) = a2;
            }
            const double hjj = Hij(j, j), hj1 = Hij(j + 1, j);
            const double beta = std::hypot(hjj, hj1);
            JFEM_CHECK(beta > 0 && std::isfinite(beta), JFEM_ERR_STATE, "FGMRES breakdown (zero Hessenberg column)");
            cs[j] = hjj / beta;
            sn[j] = hj1 / beta;
            Hij(j, j) = beta;
            Hij(j + 1, j) = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            it++;
            jn = j + 1;
            if (hist) hist[it - 1] = std::fabs(g[j + 1]) / bn;
            if (std::fabs(g[j + 1]) <= target || it >= maxit || hn == 0) stop = true;
            else if (j + 1 < m) k_scale<<<nb, KB, 0, c->stream>>>(n, 1.0 / hn, q, V.p + (size_t)(j + 1) * n);
        }
        // y = H^-1 g (upper triangular), x += Z y
        for (int i = jn - 1; i >= 0; i--) {
            double s = g[i];
            for (int k = i + 1; k < jn; k++) s -= Hij(i, k) * y[k];
            y[i] = s / Hij(i, i);
        }
        JFEM_HIP(hipMemcpyAsync(fdev, y.data(), sizeof(double) * jn, hipMemcpyHostToDevice, c->stream));
        k_lincomb<<<nb, KB, 0, c->stream>>>(n, jn, Z.p, n, fdev, x);
        JFEM_HIP(hipGetLastError());
    }
    info->iterations = (int)it;
    info->resnorm = res;
}
}  // namespace jfem

using namespace jfem;

static void check_solve_args(jfem_mat* A, jfem_mg* M, const jfem_vec* b, jfem_vec* x)
{
    JFEM_CHECK(A && b && x, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(A->rows == A->cols, JFEM_ERR_ARG, "Object is not square");
    JFEM_CHECK(b->n == A->rows && x->n == A->rows, JFEM_ERR_ARG, "Incompatible sizes");
    JFEM_CHECK(!M || mg_size(M) == A->rows, JFEM_ERR_ARG, "preconditioner has the wrong size");
}

template <typename F>
static void timed(jfem_ctx* c, jfem_solve_info* info, F&& f)
{
    std::memset(info, 0, sizeof(*info));
    JFEM_HIP(hipStreamSynchronize(c->stream));
    auto t0 = std::chrono::steady_clock::now();
    f();
    JFEM_HIP(hipStreamSynchronize(c->stream));
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

extern "C" {

int jfem_solve_cg(jfem_mat* A, jfem_mg* M, const jfem_vec* b, jfem_vec* x, double tol, long maxit,
                  jfem_solve_info* info, double* res_hist)
{
    API_BEGIN
    check_solve_args(A, M, b, x);
    jfem_solve_info tmp;
    jfem_solve_info* inf = info ? info : &tmp;
    timed(A->ctx, inf, [&] { solve_cg(A, M, b->d, x->d, tol, maxit, inf, res_hist); });
    API_END
}

int jfem_solve_fcg(jfem_mat* A, jfem_mg* M, const jfem_vec* b, jfem_vec* x, double tol, long maxit,
                   jfem_solve_info* info, double* res_hist)
{
    API_BEGIN
    check_solve_args(A, M, b, x);
    JFEM_CHECK(M, JFEM_ERR_ARG, "flexible CG needs a preconditioner");
    jfem_solve_info tmp;
    jfem_solve_info* inf = info ? info : &tmp;
    timed(A->ctx, inf, [&] { solve_fcg(A, M, b->d, x->d, tol, maxit, inf, res_hist); });
    API_END
}

int jfem_solve_bicgstab(jfem_mat* A, jfem_mg* M, const jfem_vec* b, jfem_vec* x, double tol, long maxit,
                        jfem_solve_info* info)
{
    API_BEGIN
    check_solve_args(A, M, b, x);
    jfem_solve_info tmp;
    jfem_solve_info* inf = info ? info : &tmp;
    timed(A->ctx, inf, [&] { solve_bicgstab(A, M, b->d, x->d, tol, maxit, inf); });
    API_END
}

int jfem_solve_gmres(jfem_mat* A, jfem_mg* M, const jfem_vec* b, jfem_vec* x, double tol, int restart, int max_cycles,
                     jfem_solve_info* info)
{
    API_BEGIN
    check_solve_args(A, M, b, x);
    jfem_solve_info tmp;
    jfem_solve_info* inf = info ? info : &tmp;
    timed(A->ctx, inf, [&] { solve_gmres(A, M, b->d, x->d, tol, restart, max_cycles, inf); });
    API_END
}

int jfem_smooth(jfem_mat* A, int kind, int runs, double omega, int degree, double lmin, double lmax,
                const jfem_vec* b, jfem_vec* x)
{
    API_BEGIN
    JFEM_CHECK(A && b && x, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(A->rows == A->cols && b->n == A->rows && x->n == A->rows, JFEM_ERR_ARG, "Incompatible sizes");
    JFEM_CHECK(kind == JFEM_SM_RICHARDSON || kind == JFEM_SM_JACOBI || kind == JFEM_SM_CHEBYSHEV, JFEM_ERR_UNSUPPORTED,
               "unknown smoother");
    JFEM_CHECK(runs >= 0, JFEM_ERR_ARG, "negative run count");
    if (kind == JFEM_SM_CHEBYSHEV)
        JFEM_CHECK(degree >= 1 && lmin > 0 && lmax > lmin, JFEM_ERR_ARG, "Chebyshev needs degree >= 1 and 0 < lmin < lmax");
    DevBuf<double> work(4 * (size_t)A->rows);
    smooth_any(A, kind, runs, omega, degree, lmin, lmax, b->d, x->d, work);
    API_END
}

int jfem_smooth_noalloc(jfem_mat* A, int kind, int runs, double omega, int degree, double lmin, double lmax,
                        const jfem_vec* b, jfem_vec* x)
{
    API_BEGIN
    JFEM_CHECK(A && b && x, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(A->rows == A->cols && b->n == A->rows && x->n == A->rows, JFEM_ERR_ARG, "Incompatible sizes");
    JFEM_CHECK(A->rows <= JFEM_NOALLOC_MAX_ROWS, JFEM_ERR_ARG,
               "jfem_smooth_noalloc: more than JFEM_NOALLOC_MAX_ROWS rows (its scratch holds 4 vectors of that length)");
    JFEM_CHECK(kind == JFEM_SM_RICHARDSON || kind == JFEM_SM_JACOBI || kind == JFEM_SM_CHEBYSHEV, JFEM_ERR_UNSUPPORTED,
               "unknown smoother");
    JFEM_CHECK(runs >= 0, JFEM_ERR_ARG, "negative run count");
    if (kind == JFEM_SM_CHEBYSHEV)
        JFEM_CHECK(degree >= 1 && lmin > 0 && lmax > lmin, JFEM_ERR_ARG, "Chebyshev needs degree >= 1 and 0 < lmin < lmax");
    // the device address of this GPU's copy of the static scratch array, resolved when the
    // context was created on it
    double* work = noalloc_scratch(A->ctx);
    smooth_any(A, kind, runs, omega, degree, lmin, lmax, b->d, x->d, work);
    API_END
}

int jfem_solve_gmres2(jfem_mat* A, jfem_mg* M, const jfem_vec* b, jfem_vec* x, double tol, long maxit,
                      jfem_solve_info* info)
{
    API_BEGIN
    check_solve_args(A, M, b, x);
    JFEM_CHECK(maxit >= 0, JFEM_ERR_ARG, "negative iteration limit");
    jfem_solve_info tmp;
    jfem_solve_info* inf = info ? info : &tmp;
    timed(A->ctx, inf, [&] { solve_gmres2(A, M, b->d, x->d, tol, maxit, inf); });
    API_END
}

int jfem_solve_stokes_fgmres(jfem_mat* A, jfem_mg* velocity_mg, const jfem_vec* pressure_scale,
                             const jfem_stokes_opts* opts, const jfem_vec* b, jfem_vec* x, double rtol,
                             jfem_solve_info* info, double* res_hist)
{
    API_BEGIN
    JFEM_CHECK(A && velocity_mg && pressure_scale && opts && b && x, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(A->rows == A->cols, JFEM_ERR_ARG, "Object is not square");
    JFEM_CHECK(b->n == A->rows && x->n == A->rows, JFEM_ERR_ARG, "Incompatible sizes");
    JFEM_CHECK(rtol > 0 && opts->maxit >= 0, JFEM_ERR_ARG, "bad tolerance or iteration limit");
    jfem_solve_info tmp;
    jfem_solve_info* inf = info ? info : &tmp;
    StokesPrec P(A, velocity_mg, pressure_scale->d, pressure_scale->n, opts);
    timed(A->ctx, inf, [&] { solve_stokes_fgmres(A, P, b->d, x->d, rtol, opts->restart, opts->maxit, inf, res_hist); });
    API_END
}

}  // extern "C"
