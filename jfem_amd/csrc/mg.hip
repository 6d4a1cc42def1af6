// Geometric multigrid preconditioner on nested structured Qk grids
// (MGPreconditionerInterface, M/multigrid/MGPreconditionerInterface.java:125-379;
// MGPreconditionerSpace, M/multigrid/MGPreconditionerSpace.java:28-109).
//
// Levels are re-assembled on every grid (createSystem per level) and
// Dirichlet-eliminated.  Transfers are matrix-free stencils: the reference's
// prolongation matrix P[g,f] = phi_f(x_g) of nested Qk spaces is the tensor
// product of 1D Lagrange interpolation weights, so prolongation and its
// transpose (restriction) are evaluated per node from the lattice coordinates,
// with no CSR P in HBM.  Smoothers: Richardson and Jacobi as in the reference
// (RichardsonSmoother.java:30-50, JacobiSmoother.java:28-50), and
// Chebyshev-Jacobi (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.,
// Alg. 12.1 on D^-1 A), each fused with its SpMV into one pass over the level
// matrix.  The coarsest level is solved with its dense inverse (a GEMV).
// Every level operation is templated on the working precision (FP64 or FP32).
#include "common.hpp"
#include "stream.hpp"

#include <cmath>
#include <cstring>

struct MgLevel {
    jfem::Grid grid;
    jfem_mat* A = nullptr;              // Dirichlet-eliminated level matrix (owned)
    unsigned char* bnd = nullptr;       // boundary mask (device)
    double* dinv = nullptr;             // 1/diag (1 where |diag| < 1e-10)
    float* dinv32 = nullptr;
    // work vectors, sized for doubles (reused as floats in FP32 mode)
    void *x = nullptr, *b = nullptr, *r = nullptr, *t = nullptr, *dv = nullptr;
    double lmax = 0;                    // Chebyshev upper bound on D^-1 A
};

struct jfem_mg {
    jfem_ctx* ctx = nullptr;
    std::vector<MgLevel> lev;   // 0 = coarsest
    jfem_mg_opts opt;
    double* coarse_inv = nullptr;       // dense inverse of A_0 (FP64, row-major)
    double* init = nullptr;             // FP64 scratch for the reference mvMul
    double* defect = nullptr;
    double* tmp = nullptr;
};

namespace jfem {

constexpr int MB = 256;

template <typename T>
__device__ inline T ld(const T* p, long i) { return p[i]; }

// ------------------------------------------------ fused smoother kernels
// residual of one row with L lanes: returns (A x)_row in lane 0 of the group
template <int L, typename VT, typename T>
__device__ inline T row_dot(const int* __restrict__ rowptr, const int* __restrict__ cols, const VT* __restrict__ vals,
                            const T* __restrict__ x, long row, int lane)
{
    const int s = rowptr[row], e = rowptr[row + 1];
    T acc = 0;
    for (int p = s + lane; p < e; p += L) acc += (T)vals[p] * x[cols[p]];
    for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, L);
    return acc;
}

// Jacobi: xn = x + (dinv * (b - A x)) * omega
template <int L, typename T>
__global__ __launch_bounds__(MB) void k_jacobi(long n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                               const T* __restrict__ vals, const T* __restrict__ dinv,
                                               const T* __restrict__ b, const T* __restrict__ x, T* __restrict__ xn,
                                               T omega, int x_zero)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x % L;
    const long groups = (long)gridDim.x * (MB / L);
    for (long row = (blockIdx.x * (long)MB + threadIdx.x) / L; row < n; row += groups) {
        const T ax = x_zero ? (T)0 : row_dot<L>(rowptr, cols, vals, x, row, lane);
        if (lane == 0) {
            const T xo = x_zero ? (T)0 : x[row];
            xn[row] = xo + (dinv[row] * (b[row] - ax)) * omega;
        }
    }
}

// Chebyshev step: r = b - A x; d = first ? (dinv*r)/theta : c1*d + c2*(dinv*r); xn = x + d  (c0 = theta)
template <int L, typename T>
__global__ __launch_bounds__(MB) void k_cheby(long n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                              const T* __restrict__ vals, const T* __restrict__ dinv,
                                              const T* __restrict__ b, const T* __restrict__ x, T* __restrict__ xn,
                                              T* __restrict__ d, int first, T c0, T c1, T c2, int x_zero)
{
    const int lane = threadIdx.x % L;
    const long groups = (long)gridDim.x * (MB / L);
    for (long row = (blockIdx.x * (long)MB + threadIdx.x) / L; row < n; row += groups) {
        const T ax = x_zero ? (T)0 : row_dot<L>(rowptr, cols, vals, x, row, lane);
        if (lane == 0) {
            const T z = dinv[row] * (b[row] - ax);
            const T dn = first ? z / c0 : c1 * d[row] + c2 * z;
            d[row] = dn;
            const T xo = x_zero ? (T)0 : x[row];
            xn[row] = xo + dn;
        }
    }
}

// r = b - A x
template <int L, typename T>
__global__ __launch_bounds__(MB) void k_residual(long n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                                 const T* __restrict__ vals, const T* __restrict__ b,
                                                 const T* __restrict__ x, T* __restrict__ r)
{
    const int lane = threadIdx.x % L;
    const long groups = (long)gridDim.x * (MB / L);
    for (long row = (blockIdx.x * (long)MB + threadIdx.x) / L; row < n; row += groups) {
        const T ax = row_dot<L>(rowptr, cols, vals, x, row, lane);
        if (lane == 0) r[row] = b[row] - ax;
    }
}

// ------------------------------------------------ transfers
// 1D weights of the nested Qk prolongation: fine lattice p on the fine axis,
// coarse cell C = min(p / 2k, nc - 1), position t = (p - 2kC) / 2k in C,
// weight of coarse local node l = phi_l(t) (LagrangeBasisFunction1D values).
__device__ inline double lag1(double t, int l, int k)
{
    if (k == 1) return l == 0 ? 1.0 - t : t;
    if (k == 2) {
        if (l == 0) return 1.0 - 3.0 * t + 2.0 * t * t;
        if (l == 1) return (t - t * t) * 4.0;
        return 2.0 * t * t - t;
    }
    double r = 1;
    for (int i = 0; i <= k; i++)
        if (i != l) r *= (t - 1. * i / k) / (1. * l / k - 1. * i / k);
    return r;
}

// xf = (add ? xf : 0) + P xc, with fine boundary entries zeroed if zero_bc
template <typename T>
__global__ __launch_bounds__(MB) void k_prolong(GridDev gf, GridDev gc, long nf, const T* __restrict__ xc,
                                                T* __restrict__ xf, const unsigned char* __restrict__ bndf, int add,
                                                int zero_bc)
{
    const int k = gf.k, d = gf.d;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nf; i += (long)gridDim.x * blockDim.x) {
        if (zero_bc && bndf[i]) {
            if (!add) xf[i] = 0;
            continue;
        }
        int p[MAXD];
        lattice_of_dof(gf, i, p);
        int C[MAXD];
        double w[MAXD][5];
        for (int a = 0; a < d; a++) {
            int c = p[a] / (2 * k);
            if (c > gc.n[a] - 1) c = gc.n[a] - 1;
            C[a] = c;
            const double t = (double)(p[a] - 2 * k * c) / (2 * k);
            for (int l = 0; l <= k; l++) w[a][l] = lag1(t, l, k);
        }
        int nl = 1;
        for (int a = 0; a < d; a++) nl *= k + 1;
        double acc = 0;
        for (int li = 0; li < nl; li++) {
            int r = li, q[MAXD];
            double wt = 1;
            for (int a = 0; a < d; a++) {
                const int l = r % (k + 1);
                r /= k + 1;
                wt *= w[a][l];
                q[a] = C[a] * k + l;
            }
            if (wt != 0) acc += wt * (double)xc[dof_of_lattice(gc, q)];
        }
        xf[i] = add ? (T)((double)xf[i] + acc) : (T)acc;
    }
}

// xc = P^T xf with coarse boundary entries zeroed
template <typename T>
__global__ __launch_bounds__(MB) void k_restrict(GridDev gf, GridDev gc, long nc, const T* __restrict__ xf,
                                                 T* __restrict__ xc, const unsigned char* __restrict__ bndc)
{
    const int k = gf.k, d = gf.d;
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < nc; j += (long)gridDim.x * blockDim.x) {
        if (bndc[j]) { xc[j] = 0; continue; }
        int q[MAXD];
        lattice_of_dof(gc, j, q);
        int cnt[MAXD], pf[MAXD][20];
        double w[MAXD][20];
        for (int a = 0; a < d; a++) {
            int clo, chi;
            if (q[a] % k == 0) { clo = q[a] / k - 1; chi = q[a] / k; } else { clo = chi = q[a] / k; }
            if (clo < 0) clo = 0;
            if (chi > gc.n[a] - 1) chi = gc.n[a] - 1;
            int m = 0;
            for (int p = 2 * k * clo; p <= 2 * k * (chi + 1); p++) {
                int c = p / (2 * k);
                if (c > gc.n[a] - 1) c = gc.n[a] - 1;
                const int l = q[a] - k * c;
                if (l < 0 || l > k) continue;
                const double wt = lag1((double)(p - 2 * k * c) / (2 * k), l, k);
                if (wt == 0) continue;
                pf[a][m] = p;
                w[a][m] = wt;
                m++;
            }
            cnt[a] = m;
        }
        int tot = 1;
        for (int a = 0; a < d; a++) tot *= cnt[a];
        double acc = 0;
        for (int t = 0; t < tot; t++) {
            int r = t, p[MAXD];
            double wt = 1;
            for (int a = d - 1; a >= 0; a--) {
                const int m = r % cnt[a];
                r /= cnt[a];
                p[a] = pf[a][m];
                wt *= w[a][m];
            }
            acc += wt * (double)xf[dof_of_lattice(gf, p)];
        }
        xc[j] = (T)acc;
    }
}

// x = Ainv b (dense, one wavefront per row, FP64 accumulation)
template <typename T>
__global__ __launch_bounds__(64) void k_dense_gemv(int n, const double* __restrict__ Ainv, const T* __restrict__ b,
                                                   T* __restrict__ x)
{
    const int row = blockIdx.x;
    double acc = 0;
    for (int j = threadIdx.x; j < n; j += 64) acc += Ainv[(long)row * n + j] * (double)b[j];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (threadIdx.x == 0) x[row] = (T)acc;
}

template <typename S, typename D>
__global__ void k_convert(long n, const S* __restrict__ a, D* __restrict__ b)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (D)a[i];
}

template <typename T>
__global__ void k_fill_t(long n, T v, T* __restrict__ x)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}

template <typename T>
__global__ void k_add_to(long n, const T* __restrict__ a, T* __restrict__ x, T s)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = x[i] + a[i] * s;
}

// x = mask ? v : 0  (copyBoundaryValues into a zero vector)
__global__ void k_copy_bnd(long n, const unsigned char* __restrict__ bnd, const double* __restrict__ v,
                           double* __restrict__ x)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = bnd[i] ? v[i] : 0.0;
}

// w = dinv * (A v)   (power iteration on D^-1 A)
template <int L>
__global__ __launch_bounds__(MB) void k_dinv_av(long n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                                const double* __restrict__ vals, const double* __restrict__ dinv,
                                                const double* __restrict__ v, double* __restrict__ w)
{
    const int lane = threadIdx.x % L;
    const long groups = (long)gridDim.x * (MB / L);
    for (long row = (blockIdx.x * (long)MB + threadIdx.x) / L; row < n; row += groups) {
        const double ax = row_dot<L>(rowptr, cols, vals, v, row, lane);
        if (lane == 0) w[row] = dinv[row] * ax;
    }
}

__global__ void k_dinv_from_diag(long n, const double* __restrict__ diag, double* __restrict__ dinv,
                                 float* __restrict__ dinv32)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double d = diag[i];
        if (std::fabs(d) < 1e-10) d = 1;   // JacobiSmoother ctor :20-24
        dinv[i] = 1. / d;
        dinv32[i] = (float)(1. / d);
    }
}

// deterministic pseudo-random start vector for the power iteration
__global__ void k_hash_fill(long n, double* __restrict__ v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        v[i] = 0.5 + (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
}

// ------------------------------------------------ level operations
static int mg_grid(const MgLevel& L, int lanes)
{
    const long need = (L.A->rows * lanes + MB - 1) / MB;
    const long cap = 8192;
    return (int)(need < 1 ? 1 : (need > cap ? cap : need));
}

template <typename T> static const T* vals_of(const MgLevel& L);
template <> const double* vals_of<double>(const MgLevel& L) { return L.A->vals; }
template <> const float* vals_of<float>(const MgLevel& L) { return L.A->vals32; }
template <typename T> static const T* dinv_of(const MgLevel& L);
template <> const double* dinv_of<double>(const MgLevel& L) { return L.dinv; }
template <> const float* dinv_of<float>(const MgLevel& L) { return L.dinv32; }

#define DISPATCH_L(L, ...)                              \
    switch (L) {                                        \
    case 4: { constexpr int LL = 4; __VA_ARGS__; break; }   \
    case 8: { constexpr int LL = 8; __VA_ARGS__; break; }   \
    case 16: { constexpr int LL = 16; __VA_ARGS__; break; } \
    case 32: { constexpr int LL = 32; __VA_ARGS__; break; } \
    default: { constexpr int LL = 64; __VA_ARGS__; break; } \
    }

template <typename T>
static void level_residual(jfem_mg* G, MgLevel& L, const T* b, const T* x, T* r)
{
    launch_stream<T, T>(L.A, vals_of<T>(L), x, EpiResidual<T>{b, r}, nullptr);
}

template <typename T>
static double level_norm(jfem_mg* G, long n, const T* r);

template <>
double level_norm<double>(jfem_mg* G, long n, const double* r) { return std::sqrt(vec_dot(G->ctx, n, r, r)); }
template <>
double level_norm<float>(jfem_mg* G, long n, const float* r)
{
    double* tmp = G->tmp;
    k_convert<float, double><<<grid_for(n, 256, 8192), 256, 0, G->ctx->stream>>>(n, r, tmp);
    return std::sqrt(vec_dot(G->ctx, n, tmp, tmp));
}

// Smoother.smooth(op, rhs, iterate): x (in/out, in L.x) for rhs b
template <typename T>
static void smooth(jfem_mg* G, MgLevel& L, const T* b, T*& x, T*& xalt, bool x_zero)
{
    const jfem_mg_opts& o = G->opt;
    hipStream_t st = G->ctx->stream;
    const long n = L.A->rows;
    const int g = mg_grid(L, L.A->lanes);
    if (o.smoother == JFEM_SM_JACOBI) {
        for (int it = 0; it < o.runs; it++) {
            if (x_zero && it == 0) {
                DISPATCH_L(L.A->lanes, k_jacobi<LL, T><<<g, MB, 0, st>>>(n, L.A->rowptr, L.A->colidx, vals_of<T>(L),
                                                                       dinv_of<T>(L), b, x, xalt, (T)o.omega, 1));
                JFEM_HIP(hipGetLastError());
            } else {
                launch_stream<T, T>(L.A, vals_of<T>(L), x, EpiJacobi<T>{b, x, dinv_of<T>(L), xalt, (T)o.omega}, nullptr);
            }
            std::swap(x, xalt);
        }
    } else if (o.smoother == JFEM_SM_RICHARDSON) {
        T* r = (T*)L.r;
        for (int it = 0; it < o.runs; it++) {
            if (x_zero && it == 0) k_fill_t<T><<<grid_for(n, 256, 8192), 256, 0, st>>>(n, (T)0, x);
            level_residual<T>(G, L, b, x, r);
            if (level_norm<T>(G, n, r) < 1e-10) break;   // RichardsonSmoother :36-41
            k_add_to<T><<<grid_for(n, 256, 8192), 256, 0, st>>>(n, r, x, (T)o.omega);
        }
    } else {
        const double lmax = L.lmax, lmin = L.lmax / o.cheb_ratio;
        const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
        T* dv = (T*)L.dv;
        bool zero = x_zero;
        for (int app = 0; app < o.runs; app++) {
            double rho_old = 1.0 / sigma;
            for (int kd = 0; kd < o.degree; kd++) {
                double c1 = 0, c2 = 0;
                if (kd > 0) {
                    const double rho = 1.0 / (2.0 * sigma - rho_old);
                    c1 = rho * rho_old;
                    c2 = 2.0 * rho / delta;
                    rho_old = rho;
                }
                if (zero) {
                    DISPATCH_L(L.A->lanes, k_cheby<LL, T><<<g, MB, 0, st>>>(n, L.A->rowptr, L.A->colidx, vals_of<T>(L),
                                                                          dinv_of<T>(L), b, x, xalt, dv, kd == 0 ? 1 : 0,
                                                                          (T)theta, (T)c1, (T)c2, 1));
                    JFEM_HIP(hipGetLastError());
                } else {
                    EpiCheby<T> e{b, x, dinv_of<T>(L), xalt, dv, (T)theta, (T)c1, (T)c2, kd == 0 ? 1 : 0};
                    launch_stream<T, T>(L.A, vals_of<T>(L), x, e, nullptr);
                }
                std::swap(x, xalt);
                zero = false;
            }
        }
    }
}

template <typename T>
static void restrict_to(jfem_mg* G, int lf, const T* xf, T* xc)
{
    MgLevel& F = G->lev[lf];
    MgLevel& C = G->lev[lf - 1];
    const long nc = C.A->rows;
    k_restrict<T><<<grid_for(nc, MB, 16384), MB, 0, G->ctx->stream>>>(to_dev(F.grid), to_dev(C.grid), nc, xf, xc, C.bnd);
    JFEM_HIP(hipGetLastError());
}

template <typename T>
static void prolong_to(jfem_mg* G, int lf, const T* xc, T* xf, bool add, bool zero_bc)
{
    MgLevel& F = G->lev[lf];
    MgLevel& C = G->lev[lf - 1];
    const long nf = F.A->rows;
    k_prolong<T><<<grid_for(nf, MB, 16384), MB, 0, G->ctx->stream>>>(to_dev(F.grid), to_dev(C.grid), nf, xc, xf, F.bnd,
                                                                      add ? 1 : 0, zero_bc ? 1 : 0);
    JFEM_HIP(hipGetLastError());
}

template <typename T>
static void coarse_solve(jfem_mg* G, const T* b, T* x)
{
    const int n0 = (int)G->lev[0].A->rows;
    k_dense_gemv<T><<<n0, 64, 0, G->ctx->stream>>>(n0, G->coarse_inv, b, x);
    JFEM_HIP(hipGetLastError());
}

// mgStepV(level, guess, rhs) (:125-198).  The guess is in L.x (x_zero: guess is 0),
// the rhs in L.b; the result is left in L.x.
template <typename T>
static void step_v(jfem_mg* G, int l, bool x_zero)
{
    MgLevel& L = G->lev[l];
    T* b = (T*)L.b;
    if (l == 0) { coarse_solve<T>(G, b, (T*)L.x); return; }
    T* x = (T*)L.x;
    T* xalt = (T*)L.t;
    smooth<T>(G, L, b, x, xalt, x_zero);
    MgLevel& C = G->lev[l - 1];
    T* r = (T*)L.r;
    level_residual<T>(G, L, b, x, r);
    restrict_to<T>(G, l, r, (T*)C.b);
    step_v<T>(G, l - 1, true);
    prolong_to<T>(G, l, (const T*)C.x, x, true, true);   // guess.add(P e) with zero BC
    smooth<T>(G, L, b, x, xalt, false);
    // keep the result in L.x / the spare in L.t
    if (x != (T*)L.x) { L.t = L.x; L.x = x; }
}

template <typename T>
static void apply_t(jfem_mg* G, const double* rin, double* zout)
{
    const int Lf = (int)G->lev.size() - 1;
    MgLevel& F = G->lev[Lf];
    const long n = F.A->rows;
    hipStream_t st = G->ctx->stream;
    const int g = grid_for(n, 256, 16384);
    if (!G->opt.fmg) {
        k_convert<double, T><<<g, 256, 0, st>>>(n, rin, (T*)F.b);
        for (int c = 0; c < G->opt.cycles; c++) step_v<T>(G, Lf, c == 0);
        k_convert<T, double><<<g, 256, 0, st>>>(n, (const T*)F.x, zout);
        JFEM_HIP(hipGetLastError());
        return;
    }
    // MGPreconditionerInterface.mvMul (:352-379): initial = copyBoundaryValues(v),
    // defect = v - A initial, full-V correction, cycles V-cycles, initial + iterate
    k_copy_bnd<<<g, 256, 0, st>>>(n, F.bnd, rin, G->init);
    spmv(F.A, G->init, G->tmp);
    vec_copy(G->ctx, G->defect, rin, n);
    vec_axpby(G->ctx, n, -1.0, G->tmp, 1.0, G->defect);
    // fullVCycleCorrection (:307-326): restrict the defect to every level
    k_convert<double, T><<<g, 256, 0, st>>>(n, G->defect, (T*)F.r);
    const T* src = (const T*)F.r;
    for (int l = Lf; l >= 1; l--) {
        restrict_to<T>(G, l, src, (T*)G->lev[l - 1].r);
        src = (const T*)G->lev[l - 1].r;
    }
    for (int l = 0; l < Lf; l++) {
        MgLevel& L = G->lev[l];
        JFEM_HIP(hipMemcpyAsync(L.b, L.r, sizeof(T) * L.A->rows, hipMemcpyDeviceToDevice, st));
        step_v<T>(G, l, l == 0);
        prolong_to<T>(G, l + 1, (const T*)L.x, (T*)G->lev[l + 1].x, false, false);   // raw prolongation matrix
    }
    k_convert<double, T><<<g, 256, 0, st>>>(n, G->defect, (T*)F.b);
    for (int c = 0; c < G->opt.cycles; c++) step_v<T>(G, Lf, false);
    k_convert<T, double><<<g, 256, 0, st>>>(n, (const T*)F.x, zout);
    vec_axpby(G->ctx, n, 1.0, G->init, 1.0, zout);
    JFEM_HIP(hipGetLastError());
}

void mg_apply(jfem_mg* G, const double* r, double* z)
{
    if (G->opt.precision == JFEM_FP32) apply_t<float>(G, r, z);
    else apply_t<double>(G, r, z);
}

long mg_size(const jfem_mg* G) { return G->lev.back().A->rows; }

// ------------------------------------------------ setup
// textbook LU with partial pivoting (Golub & Van Loan 3.4.1) and the explicit
// inverse of the coarsest matrix
static std::vector<double> dense_inverse(int n, std::vector<double> A)
{
    std::vector<int> piv(n);
    for (int i = 0; i < n; i++) piv[i] = i;
    for (int k = 0; k < n; k++) {
        int pr = k;
        double best = std::fabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (std::fabs(A[(size_t)i * n + k]) > best) { best = std::fabs(A[(size_t)i * n + k]); pr = i; }
        JFEM_CHECK(best != 0, JFEM_ERR_STATE, "coarse matrix is singular");
        if (pr != k) {
            for (int j = 0; j < n; j++) std::swap(A[(size_t)k * n + j], A[(size_t)pr * n + j]);
            std::swap(piv[k], piv[pr]);
        }
        for (int i = k + 1; i < n; i++) {
            const double l = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            A[(size_t)i * n + k] = l;
            for (int j = k + 1; j < n; j++) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
        }
    }
    std::vector<double> inv((size_t)n * n), y(n);
    for (int col = 0; col < n; col++) {
        for (int i = 0; i < n; i++) {
            double s = piv[i] == col ? 1.0 : 0.0;
            for (int j = 0; j < i; j++) s -= A[(size_t)i * n + j] * y[j];
            y[i] = s;
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = y[i];
            for (int j = i + 1; j < n; j++) s -= A[(size_t)i * n + j] * y[j];
            y[i] = s / A[(size_t)i * n + i];
        }
        for (int i = 0; i < n; i++) inv[(size_t)i * n + col] = y[i];
    }
    return inv;
}

static double power_lmax(jfem_mg* G, MgLevel& L, int its)
{
    jfem_ctx* c = G->ctx;
    const long n = L.A->rows;
    double* v = (double*)L.r;
    double* w = (double*)L.t;
    hipStream_t st = c->stream;
    k_hash_fill<<<grid_for(n, 256, 8192), 256, 0, st>>>(n, v);
    double nv = std::sqrt(vec_dot(c, n, v, v));
    vec_axpby(c, n, 0.0, v, 1.0 / nv, v);
    double lam = 0;
    const int g = mg_grid(L, L.A->lanes);
    for (int i = 0; i < its; i++) {
        DISPATCH_L(L.A->lanes, k_dinv_av<LL><<<g, MB, 0, st>>>(n, L.A->rowptr, L.A->colidx, L.A->vals, L.dinv, v, w));
        JFEM_HIP(hipGetLastError());
        const double nw = std::sqrt(vec_dot(c, n, w, w));
        lam = nw;
        if (nw == 0) break;
        vec_axpby(c, n, 1.0 / nw, w, 0.0, v);
    }
    return lam;
}

static void free_mg(jfem_mg* G)
{
    if (!G) return;
    (void)hipStreamSynchronize(G->ctx->stream);
    for (auto& L : G->lev) {
        mat_free(L.A);
        (void)hipFree(L.bnd); (void)hipFree(L.dinv); (void)hipFree(L.dinv32);
        (void)hipFree(L.x); (void)hipFree(L.b); (void)hipFree(L.r); (void)hipFree(L.t); (void)hipFree(L.dv);
    }
    (void)hipFree(G->coarse_inv);
    (void)hipFree(G->init); (void)hipFree(G->defect); (void)hipFree(G->tmp);
    delete G;
}

static jfem_mg* create_mg(jfem_space* const* spaces, int nlev, int nint, const int* kinds, const double* weights,
                          const double* coeff, int rule, const jfem_mg_opts* opts)
{
    JFEM_CHECK(spaces && nlev >= 1 && opts, JFEM_ERR_ARG, "null argument");
    jfem_ctx* c = spaces[0]->ctx;
    for (int l = 1; l < nlev; l++) {
        const Grid& a = spaces[l - 1]->grid;
        const Grid& b = spaces[l]->grid;
        JFEM_CHECK(a.d == b.d && a.k == b.k, JFEM_ERR_ARG, "levels differ in dimension or degree");
        for (int i = 0; i < a.d; i++) {
            JFEM_CHECK(b.n[i] == 2 * a.n[i], JFEM_ERR_ARG, "each level must double the cells of the previous one");
            JFEM_CHECK(a.start[i] == b.start[i] && a.end[i] == b.end[i], JFEM_ERR_ARG, "levels cover different boxes");
        }
    }
    JFEM_CHECK(opts->smoother >= 0 && opts->smoother <= 2, JFEM_ERR_ARG, "unknown smoother");
    JFEM_CHECK(opts->cycles >= 1, JFEM_ERR_ARG, "cycles must be positive");
    JFEM_CHECK(opts->smoother != JFEM_SM_CHEBYSHEV || (opts->degree >= 1 && opts->cheb_ratio > 1), JFEM_ERR_ARG,
               "bad Chebyshev parameters");
    JFEM_CHECK(spaces[0]->grid.ndofs() <= 8192, JFEM_ERR_ARG, "coarsest level too large for the dense solve");
    auto* G = new jfem_mg();
    G->ctx = c;
    G->opt = *opts;
    G->lev.resize(nlev);
    try {
        // per-level coefficients: the finest given, coarser = mean of the 2^d children
        std::vector<std::vector<double>> coef(nlev);
        if (coeff) {
            const Grid& gf = spaces[nlev - 1]->grid;
            coef[nlev - 1].assign(coeff, coeff + gf.ncells());
            for (int l = nlev - 2; l >= 0; l--) {
                const Grid& gc = spaces[l]->grid;
                const Grid& gq = spaces[l + 1]->grid;
                coef[l].assign(gc.ncells(), 0.0);
                for (long cc = 0; cc < gc.ncells(); cc++) {
                    int ci[MAXD] = {0, 0, 0};
                    long r = cc;
                    for (int a = gc.d - 1; a >= 0; a--) { ci[a] = (int)(r % gc.n[a]); r /= gc.n[a]; }
                    double s = 0;
                    for (int ch = 0; ch < (1 << gc.d); ch++) {
                        long fl = 0;
                        for (int a = 0; a < gc.d; a++) fl = fl * gq.n[a] + 2 * ci[a] + ((ch >> a) & 1);
                        s += coef[l + 1][fl];
                    }
                    coef[l][cc] = s / (1 << gc.d);
                }
            }
        }
        for (int l = 0; l < nlev; l++) {
            MgLevel& L = G->lev[l];
            L.grid = spaces[l]->grid;
            L.A = assemble_matrix(spaces[l], nint, kinds, weights, coeff ? coef[l].data() : nullptr, rule);
            const long n = L.A->rows;
            JFEM_HIP(hipMalloc(&L.bnd, n));
            space_boundary_mask(spaces[l], L.bnd);
            apply_dirichlet_mask(L.A, nullptr, L.bnd, nullptr);
            if (opts->precision == JFEM_FP32) mat_make_f32(L.A);
            JFEM_HIP(hipMalloc(&L.dinv, sizeof(double) * n));
            JFEM_HIP(hipMalloc(&L.dinv32, sizeof(float) * n));
            for (void** p : {&L.x, &L.b, &L.r, &L.t, &L.dv}) {
                JFEM_HIP(hipMalloc(p, sizeof(double) * n));
                JFEM_HIP(hipMemsetAsync(*p, 0, sizeof(double) * n, c->stream));
            }
            mat_diag(L.A, (double*)L.t);
            k_dinv_from_diag<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, (double*)L.t, L.dinv, L.dinv32);
            JFEM_HIP(hipGetLastError());
            if (opts->smoother == JFEM_SM_CHEBYSHEV && l > 0)
                L.lmax = (opts->cheb_safety > 0 ? opts->cheb_safety : 1.1) *
                         power_lmax(G, L, opts->power_its > 0 ? opts->power_its : 30);
        }
        // coarsest dense inverse
        MgLevel& C = G->lev[0];
        const int n0 = (int)C.A->rows;
        std::vector<int> rp(n0 + 1), ci(C.A->nnz);
        std::vector<double> cv(C.A->nnz), dense((size_t)n0 * n0, 0.0);
        JFEM_HIP(hipMemcpy(rp.data(), C.A->rowptr, sizeof(int) * (n0 + 1), hipMemcpyDeviceToHost));
        JFEM_HIP(hipMemcpy(ci.data(), C.A->colidx, sizeof(int) * C.A->nnz, hipMemcpyDeviceToHost));
        JFEM_HIP(hipMemcpy(cv.data(), C.A->vals, sizeof(double) * C.A->nnz, hipMemcpyDeviceToHost));
        for (int i = 0; i < n0; i++)
            for (int p = rp[i]; p < rp[i + 1]; p++) dense[(size_t)i * n0 + ci[p]] += cv[p];
        std::vector<double> inv = dense_inverse(n0, dense);
        JFEM_HIP(hipMalloc(&G->coarse_inv, sizeof(double) * inv.size()));
        JFEM_HIP(hipMemcpy(G->coarse_inv, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice));
        const long nf = G->lev.back().A->rows;
        JFEM_HIP(hipMalloc(&G->init, sizeof(double) * nf));
        JFEM_HIP(hipMalloc(&G->defect, sizeof(double) * nf));
        JFEM_HIP(hipMalloc(&G->tmp, sizeof(double) * nf));
        JFEM_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
        free_mg(G);
        throw;
    }
    return G;
}

}  // namespace jfem

using namespace jfem;

extern "C" {

int jfem_mg_create(jfem_space* const* spaces, int nlev, int nint, const int* kinds, const double* weights,
                   const double* cell_coeff, int quad_rule, const jfem_mg_opts* opts, jfem_mg** out)
{
    API_BEGIN
    JFEM_CHECK(out && kinds && weights, JFEM_ERR_ARG, "null argument");
    *out = create_mg(spaces, nlev, nint, kinds, weights, cell_coeff, quad_rule, opts);
    API_END
}

int jfem_mg_destroy(jfem_mg* mg)
{
    API_BEGIN
    free_mg(mg);
    API_END
}

int jfem_mg_level_matrix(jfem_mg* mg, int level, jfem_mat** out)
{
    API_BEGIN
    JFEM_CHECK(mg && out, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(level >= 0 && level < (int)mg->lev.size(), JFEM_ERR_ARG, "no such level");
    *out = mg->lev[level].A;
    API_END
}

int jfem_mg_apply(jfem_mg* mg, const jfem_vec* r, jfem_vec* z)
{
    API_BEGIN
    JFEM_CHECK(mg && r && z, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(r->n == mg_size(mg) && z->n == mg_size(mg), JFEM_ERR_ARG, "defect has wrong shape");
    JFEM_CHECK(r->d != z->d, JFEM_ERR_ARG, "output aliases input");
    mg_apply(mg, r->d, z->d);
    JFEM_HIP(hipStreamSynchronize(mg->ctx->stream));
    API_END
}

int jfem_mg_vcycle(jfem_mg* mg, jfem_vec* x, const jfem_vec* b)
{
    API_BEGIN
    JFEM_CHECK(mg && x && b, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(x->n == mg_size(mg) && b->n == mg_size(mg), JFEM_ERR_ARG, "wrong size");
    const int Lf = (int)mg->lev.size() - 1;
    MgLevel& F = mg->lev[Lf];
    const long n = F.A->rows;
    hipStream_t st = mg->ctx->stream;
    const int g = grid_for(n, 256, 16384);
    if (mg->opt.precision == JFEM_FP32) {
        k_convert<double, float><<<g, 256, 0, st>>>(n, b->d, (float*)F.b);
        k_convert<double, float><<<g, 256, 0, st>>>(n, x->d, (float*)F.x);
        step_v<float>(mg, Lf, false);
        k_convert<float, double><<<g, 256, 0, st>>>(n, (const float*)F.x, x->d);
    } else {
        vec_copy(mg->ctx, (double*)F.b, b->d, n);
        vec_copy(mg->ctx, (double*)F.x, x->d, n);
        step_v<double>(mg, Lf, false);
        vec_copy(mg->ctx, x->d, (const double*)F.x, n);
    }
    JFEM_HIP(hipStreamSynchronize(st));
    API_END
}

int jfem_mg_lmax(jfem_mg* mg, double* out)
{
    API_BEGIN
    JFEM_CHECK(mg && out, JFEM_ERR_ARG, "null argument");
    for (size_t l = 0; l < mg->lev.size(); l++) out[l] = mg->lev[l].lmax;
    API_END
}

}  // extern "C"
