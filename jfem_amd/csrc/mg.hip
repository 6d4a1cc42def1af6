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
#include "nn.hpp"
#include "stencil.hpp"
#include "stream.hpp"

#include <chrono>
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
    // transfer from the next coarser level (levels > 0): per axis the 1D
    // prolongation P_a (fine lattice x coarse lattice) and R_a = P_a^T as CSR
    // in one device buffer, and two lattice-ordered work arrays
    int* tp_rp[2][jfem::MAXD] = {};           // [0] = P, [1] = R: row pointers
    int* tp_ix[2][jfem::MAXD] = {};           // column (lattice) indices
    double* tp_w[2][jfem::MAXD] = {};         // weights
    void* tp_buf = nullptr;             // owns the tables above
    void *tw0 = nullptr, *tw1 = nullptr;
    // matrix-free operator (smoothing and residuals without reading A): parameters and the
    // level's cell coefficients (nullptr: constant 1)
    bool mf = false;
    jfem::MfParams mfp;
    double* coef = nullptr;
    float* coef32 = nullptr;
    unsigned short* mf_cand = nullptr;   // the inner-tile table of mfp (device)
};

// learned smoother on the finest level (a network applied to the normalised
// residual on the node lattice, x += N(r / rms(r)) * rms(r))
struct LearnedState {
    jfem_nn* nn = nullptr;       // borrowed
    jfem::Lat lat;               // finest node lattice
    float* stat = nullptr;       // np x (cin - 1) static input channels
    int* perm = nullptr;         // lattice index of every finest DOF
    float* stat_dof = nullptr;   // the static channels in DOF order
    float* slog = nullptr;       // the log-coefficient channel in lattice order (nullptr: zero)
    int* ip = nullptr;           // DOF of every lattice pixel, -1 on Dirichlet nodes
    int runs = 1;                // learned steps per smoothing call
    int sides = 3;               // smoothing calls that use the network: 1 pre-smoothing, 2 post-smoothing, 3 both
    int active = 0;
    int fell_back = 0;
    double fallback = 0.5;       // switch to the plain smoother when a cycle reduces less than this
    double* rr = nullptr;        // device scalar: squared residual norm
    double *u = nullptr, *z = nullptr, *gd = nullptr, *rh = nullptr, *xs = nullptr, *bs = nullptr;   // training work
    double* loss = nullptr;      // device loss history
    long loss_cap = 0;
    int mr = 0;                  // step rule: 0 plain x += N(r); 1 minimal residual; 2 minimal energy along N(r)
    int capture = 0;             // training: record the residual before finest post-smoothing
    double* capbuf = nullptr;
    double* capx = nullptr;      // training: the iterate before finest post-smoothing
    double *xt = nullptr, *eh = nullptr, *w = nullptr, *y = nullptr;   // two-grid training work
    double* loss2 = nullptr;     // two-grid training: (numerator, denominator) per step
    long loss2_cap = 0;
};

struct jfem_mg {
    jfem_ctx* ctx = nullptr;
    std::vector<MgLevel> lev;   // 0 = coarsest
    jfem_mg_opts opt;
    double* coarse_inv = nullptr;       // dense inverse of A_0 (FP64, row-major)
    double* init = nullptr;             // FP64 scratch for the reference mvMul
    double* defect = nullptr;
    double* tmp = nullptr;
    LearnedState ls;
};

namespace jfem {

constexpr int MB = 256;

// ------------------------------------------------ smoother kernels
// first Jacobi step from x = 0: xn = x + (dinv * (b - A x)) * omega with x = 0
// (JacobiSmoother.java:45-48); no matrix is read
template <typename T>
__global__ __launch_bounds__(MB) void k_jacobi0(long n, const T* __restrict__ dinv, const T* __restrict__ b,
                                                T* __restrict__ xn, T omega)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)MB + threadIdx.x; i < n; i += (long)gridDim.x * MB)
        xn[i] = (T)0 + (dinv[i] * (b[i] - (T)0)) * omega;
}

// first Chebyshev step from x = 0: d = dinv * (b - A x) / theta, xn = x + d with x = 0
template <typename T>
__global__ __launch_bounds__(MB) void k_cheby0(long n, const T* __restrict__ dinv, const T* __restrict__ b,
                                               T* __restrict__ xn, T* __restrict__ d, T theta)
{
#pragma clang fp contract(off)
    for (long i = blockIdx.x * (long)MB + threadIdx.x; i < n; i += (long)gridDim.x * MB) {
        const T dn = (dinv[i] * (b[i] - (T)0)) / theta;
        d[i] = dn;
        xn[i] = (T)0 + dn;
    }
}

// ------------------------------------------------ transfers
// The prolongation of nested tensor-product Qk spaces is a Kronecker product
// P = P_{d-1} x ... x P_0 of 1D matrices: fine lattice node p on an axis lies in
// coarse cell C = min(p / 2k, nc - 1) at t = (p - 2kC) / 2k, and its weight for
// coarse local node l is the Lagrange value phi_l(t) (LagrangeBasisFunction1D
// values).  P and P^T are applied one axis at a time (sum factorisation): each
// pass is a 1D sparse matrix on one lattice axis, the first pass reads the
// source in DOF order and the last writes the target in DOF order, the passes
// between work on lattice-ordered arrays (axis 0 fastest).
__host__ __device__ inline double lag1(double t, int l, int k)
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

// dof_of_lattice / lattice_of_dof with D fixed, in 32-bit arithmetic: the levels of a hierarchy
// have fewer than 2^31 nodes (checked when a level is built), so every partial product and every
// index fits, and the divisions are 32-bit (a 64-bit division is an emulated sequence)
template <int D, int KK>
__device__ inline long dof_of_lat_k(const GridDev& g, const int* p)
{
    const unsigned k = KK > 0 ? (unsigned)KK : (unsigned)g.k;
    unsigned tail[D];
    unsigned t = 1;
#pragma unroll
    for (int a = D - 1; a >= 0; a--) { tail[a] = t; t *= k * (unsigned)g.n[a] + 1u; }
    unsigned pre = 0, mprod = 1, rank = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const unsigned pa = (unsigned)p[a];
        const unsigned c = pa == 0 ? 0u : (pa - 1u) / k;
        const unsigned l = pa - c * k;
        const unsigned m = c > 0 ? k : k + 1u;
        const unsigned S = c > 0 ? c * k + 1u : 0u;
        pre += mprod * S * tail[a];
        rank += (l - (c > 0 ? 1u : 0u)) * mprod;
        mprod *= m;
    }
    return (long)(pre + rank);
}

template <int D, int KK>
__device__ inline void lat_of_dof_k(const GridDev& g, long dof, int* p)
{
    const unsigned k = KK > 0 ? (unsigned)KK : (unsigned)g.k;
    unsigned tail[D];
    unsigned t = 1;
#pragma unroll
    for (int a = D - 1; a >= 0; a--) { tail[a] = t; t *= k * (unsigned)g.n[a] + 1u; }
    unsigned rem = (unsigned)dof, mprod = 1;
    unsigned c[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        const unsigned unit = mprod * tail[a];
        const unsigned q = rem / unit;
        c[a] = q < k + 1u ? 0u : (q - 1u) / k;
        rem -= (c[a] > 0 ? c[a] * k + 1u : 0u) * unit;
        mprod *= c[a] > 0 ? k : k + 1u;
    }
#pragma unroll
    for (int a = 0; a < D; a++) {
        const unsigned m = c[a] > 0 ? k : k + 1u;
        p[a] = (int)(c[a] * k + rem % m + (c[a] > 0 ? 1u : 0u));
        rem /= m;
    }
}

// the degree-2 grids of the Q2 hierarchies with k fixed at compile time (divisions by 2 and 3)
template <int D>
__device__ inline long dof_of_lat(const GridDev& g, const int* p)
{
    return g.k == 2 ? dof_of_lat_k<D, 2>(g, p) : dof_of_lat_k<D, 0>(g, p);
}

template <int D>
__device__ inline void lat_of_dof(const GridDev& g, long dof, int* p)
{
    if (g.k == 2) lat_of_dof_k<D, 2>(g, dof, p);
    else lat_of_dof_k<D, 0>(g, dof, p);
}

struct Shape3 {
    int n[3];
};

// One pass: out = M in along lattice axis AX (M = 1D CSR rp/ix/w).  IN_DOF: the
// source is a DOF-ordered vector of grid gin; otherwise a lattice array of
// shape sin.  OUT_DOF: the target is a DOF-ordered vector of grid gout with
// one thread per DOF (coalesced stores), entries with bnd set are left alone
// (add) or zeroed; otherwise a lattice array of shape sout.  Sums in FP64.
template <typename T, int D, int AX, bool IN_DOF, bool OUT_DOF>
__global__ __launch_bounds__(MB) void k_tp_pass(Shape3 sin, Shape3 sout, GridDev gin, GridDev gout,
                                                const int* __restrict__ rp, const int* __restrict__ ix,
                                                const double* __restrict__ w, const T* __restrict__ in,
                                                T* __restrict__ out, const unsigned char* __restrict__ bnd, int add)
{
    const long total = (long)sout.n[0] * sout.n[1] * sout.n[2];
    for (long t = blockIdx.x * (long)MB + threadIdx.x; t < total; t += (long)gridDim.x * MB) {
        int q[3] = {0, 0, 0};
        if (OUT_DOF) {
            if (bnd && bnd[t]) {
                if (!add) out[t] = (T)0;
                continue;
            }
            lat_of_dof<D>(gout, t, q);
        } else {
            // lattice arrays keep the last axis fastest, the fast axis of the DOF order: the first
            // and last passes then gather or scatter DOF-ordered data along short runs
            q[2] = (int)(t % sout.n[2]);
            const long r = t / sout.n[2];
            q[1] = (int)(r % sout.n[1]);
            q[0] = (int)(r / sout.n[1]);
        }
        int p[3] = {q[0], q[1], q[2]};
        const int row = q[AX];
        double acc = 0;
        for (int e = rp[row]; e < rp[row + 1]; e++) {
            p[AX] = ix[e];
            const long src = IN_DOF ? dof_of_lat<D>(gin, p) : p[2] + (long)sin.n[2] * (p[1] + (long)sin.n[1] * p[0]);
            acc += w[e] * (double)in[src];
        }
        if (OUT_DOF && add) acc += (double)out[t];
        out[t] = (T)acc;
    }
}

template <typename T, int D, int AX>
static void tp_pass_ax(hipStream_t st, bool in_dof, bool out_dof, Shape3 sin, Shape3 sout, GridDev gin, GridDev gout,
                       const int* rp, const int* ix, const double* w, const T* in, T* out, const unsigned char* bnd,
                       int add)
{
    const long total = (long)sout.n[0] * sout.n[1] * sout.n[2];
    const int g = grid_for(total, MB, 65536);
    if (in_dof && out_dof)
        k_tp_pass<T, D, AX, true, true><<<g, MB, 0, st>>>(sin, sout, gin, gout, rp, ix, w, in, out, bnd, add);
    else if (in_dof)
        k_tp_pass<T, D, AX, true, false><<<g, MB, 0, st>>>(sin, sout, gin, gout, rp, ix, w, in, out, bnd, add);
    else if (out_dof)
        k_tp_pass<T, D, AX, false, true><<<g, MB, 0, st>>>(sin, sout, gin, gout, rp, ix, w, in, out, bnd, add);
    else
        k_tp_pass<T, D, AX, false, false><<<g, MB, 0, st>>>(sin, sout, gin, gout, rp, ix, w, in, out, bnd, add);
    JFEM_HIP(hipGetLastError());
}

template <typename T, int D>
static void tp_pass_d(hipStream_t st, int ax, bool in_dof, bool out_dof, Shape3 sin, Shape3 sout, GridDev gin,
                      GridDev gout, const int* rp, const int* ix, const double* w, const T* in, T* out,
                      const unsigned char* bnd, int add)
{
    if (ax == 0) tp_pass_ax<T, D, 0>(st, in_dof, out_dof, sin, sout, gin, gout, rp, ix, w, in, out, bnd, add);
    else if (ax == 1) tp_pass_ax<T, D, (D > 1 ? 1 : 0)>(st, in_dof, out_dof, sin, sout, gin, gout, rp, ix, w, in, out, bnd, add);
    else tp_pass_ax<T, D, (D > 2 ? 2 : 0)>(st, in_dof, out_dof, sin, sout, gin, gout, rp, ix, w, in, out, bnd, add);
}

// Apply P (restrict = false: xf = (add ? xf : 0) + P xc, fine entries with bnd
// set left alone / zeroed) or P^T (restrict = true: xc = P^T xf, coarse entries
// with bnd set zeroed) of level F (fine) over level C (coarse).
template <typename T>
static void tp_transfer(hipStream_t st, const MgLevel& F, const MgLevel& C, bool restrict_, const T* src, T* dst,
                        const unsigned char* bnd, int add)
{
    const int d = F.grid.d;
    const GridDev gf = to_dev(F.grid), gc = to_dev(C.grid);
    Shape3 cur;    // lattice shape of the current source
    for (int a = 0; a < 3; a++) cur.n[a] = a < d ? (int)(restrict_ ? F.grid.nodes_axis(a) : C.grid.nodes_axis(a)) : 1;
    const T* in = src;
    for (int s = 0; s < d; s++) {
        const int ax = restrict_ ? s : d - 1 - s;
        const bool first = s == 0, last = s == d - 1;
        Shape3 nxt = cur;
        nxt.n[ax] = (int)(restrict_ ? C.grid.nodes_axis(ax) : F.grid.nodes_axis(ax));
        // intermediate shapes: restriction (Nc0, Nf1, Nf2) -> tw0, (Nc0, Nc1, Nf2) -> tw1;
        // prolongation (Nc0, Nc1, Nf2) -> tw1, (Nc0, Nf1, Nf2) -> tw0
        const bool big = restrict_ ? s == 0 : s == d - 2;
        T* out = last ? dst : (T*)(big ? F.tw0 : F.tw1);
        const int which = restrict_ ? 1 : 0;
        const GridDev gin = restrict_ ? gf : gc, gout = restrict_ ? gc : gf;
        const unsigned char* b = last ? bnd : nullptr;
        const int ad = last ? add : 0;
        switch (d) {
        case 1: tp_pass_d<T, 1>(st, ax, first, last, cur, nxt, gin, gout, F.tp_rp[which][ax], F.tp_ix[which][ax], F.tp_w[which][ax], in, out, b, ad); break;
        case 2: tp_pass_d<T, 2>(st, ax, first, last, cur, nxt, gin, gout, F.tp_rp[which][ax], F.tp_ix[which][ax], F.tp_w[which][ax], in, out, b, ad); break;
        default: tp_pass_d<T, 3>(st, ax, first, last, cur, nxt, gin, gout, F.tp_rp[which][ax], F.tp_ix[which][ax], F.tp_w[which][ax], in, out, b, ad); break;
        }
        in = out;
        cur = nxt;
    }
}

// 1D tables of the transfer onto fine level F from coarse level C, and the work arrays
static void build_transfer(MgLevel& F, const MgLevel& C)
{
    const int d = F.grid.d, k = F.grid.k;
    std::vector<int> ints;
    std::vector<double> dbl;
    struct Ofs { size_t rp, ix, w; };
    Ofs ofs[2][MAXD];
    for (int a = 0; a < d; a++) {
        const int nc = C.grid.n[a];
        const int Nf = (int)F.grid.nodes_axis(a), Nc = (int)C.grid.nodes_axis(a);
        // P: row p (fine), entries (coarse node, weight) with nonzero weight
        std::vector<std::vector<std::pair<int, double>>> P(Nf), R(Nc);
        for (int p = 0; p < Nf; p++) {
            int c = p / (2 * k);
            if (c > nc - 1) c = nc - 1;
            const double t = (double)(p - 2 * k * c) / (2 * k);
            for (int l = 0; l <= k; l++) {
                const double wt = lag1(t, l, k);
                if (wt != 0) P[p].push_back({c * k + l, wt});
            }
        }
        for (int p = 0; p < Nf; p++)
            for (auto& e : P[p]) R[e.first].push_back({p, e.second});
        for (int m = 0; m < 2; m++) {
            const auto& M = m == 0 ? P : R;
            ofs[m][a].rp = ints.size();
            int cnt = 0;
            ints.push_back(0);
            for (auto& row : M) { cnt += (int)row.size(); ints.push_back(cnt); }
            ofs[m][a].ix = ints.size();
            ofs[m][a].w = dbl.size();
            for (auto& row : M)
                for (auto& e : row) { ints.push_back(e.first); dbl.push_back(e.second); }
        }
    }
    const size_t ib = sizeof(int) * ints.size(), db = sizeof(double) * dbl.size();
    const size_t ioff = (db + 255) / 256 * 256;
    char* buf = nullptr;
    JFEM_HIP(hipMalloc(&buf, ioff + ib));
    F.tp_buf = buf;
    JFEM_HIP(hipMemcpy(buf, dbl.data(), db, hipMemcpyHostToDevice));
    JFEM_HIP(hipMemcpy(buf + ioff, ints.data(), ib, hipMemcpyHostToDevice));
    for (int m = 0; m < 2; m++)
        for (int a = 0; a < d; a++) {
            F.tp_rp[m][a] = (int*)(buf + ioff) + ofs[m][a].rp;
            F.tp_ix[m][a] = (int*)(buf + ioff) + ofs[m][a].ix;
            F.tp_w[m][a] = (double*)buf + ofs[m][a].w;
        }
    // work arrays: restriction writes (Nc0, Nf1, Nf2) then (Nc0, Nc1, Nf2);
    // prolongation writes (Nc0, Nc1, Nf2) then (Nc0, Nf1, Nf2)
    long n0 = 1, n1 = 1;
    for (int a = 0; a < d; a++) {
        const long Nf = F.grid.nodes_axis(a), Nc = C.grid.nodes_axis(a);
        n0 *= a == 0 ? Nc : Nf;
        n1 *= a <= 1 ? Nc : Nf;
    }
    JFEM_HIP(hipMalloc(&F.tw0, sizeof(double) * n0));
    JFEM_HIP(hipMalloc(&F.tw1, sizeof(double) * n1));
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

// w = dinv * (A v): stream epilogue of the power iteration on D^-1 A
struct EpiDinv {
    static constexpr bool reduce = false;
    const double* dinv;
    double* w;
    struct Pre { double dinv; };
    __device__ Pre load(long row) const { return {dinv[row]}; }
    __device__ double operator()(long row, double ax, const Pre& p) const { w[row] = p.dinv * ax; return 0.0; }
};

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
template <typename T> static const T* vals_of(const MgLevel& L);
template <> const double* vals_of<double>(const MgLevel& L) { return L.A->vals; }
template <> const float* vals_of<float>(const MgLevel& L) { return L.A->vals32; }
template <typename T> static const T* dinv_of(const MgLevel& L);
template <> const double* dinv_of<double>(const MgLevel& L) { return L.dinv; }
template <> const float* dinv_of<float>(const MgLevel& L) { return L.dinv32; }

template <typename T> static const T* coef_of(const MgLevel& L);
template <> const double* coef_of<double>(const MgLevel& L) { return L.coef; }
template <> const float* coef_of<float>(const MgLevel& L) { return L.coef32; }

// (A_l x)_row handed to a row epilogue: matrix-free on levels that have the operator,
// otherwise the CSR-stream product with the level matrix
template <typename T, typename Epi>
static void level_apply(jfem_mg* G, const MgLevel& L, const T* x, const Epi& epi)
{
    if (L.mf && launch_mf<T>(G->ctx->stream, L.mfp, x, coef_of<T>(L), epi)) return;
    launch_stream<T, T>(L.A, vals_of<T>(L), x, epi, nullptr);
}

template <typename T>
static void level_residual(jfem_mg* G, MgLevel& L, const T* b, const T* x, T* r)
{
    level_apply<T>(G, L, x, EpiResidual<T>{b, r});
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

template <typename T>
static void learned_smooth(jfem_mg* G, MgLevel& L, const T* b, T* x, T* u, bool x_zero);

// Smoother.smooth(op, rhs, iterate): x (in/out, in L.x) for rhs b; post marks
// the post-smoothing call of a V-cycle
template <typename T>
static void smooth(jfem_mg* G, MgLevel& L, const T* b, T*& x, T*& xalt, bool x_zero, bool post)
{
    const jfem_mg_opts& o = G->opt;
    hipStream_t st = G->ctx->stream;
    const long n = L.A->rows;
    if (G->ls.active && &L == &G->lev.back() && (G->ls.sides & (post ? 2 : 1))) {
        learned_smooth<T>(G, L, b, x, xalt, x_zero);
        return;
    }
    if (o.smoother == JFEM_SM_JACOBI) {
        for (int it = 0; it < o.runs; it++) {
            if (x_zero && it == 0) {
                k_jacobi0<T><<<grid_for(n, MB, 16384), MB, 0, st>>>(n, dinv_of<T>(L), b, xalt, (T)o.omega);
                JFEM_HIP(hipGetLastError());
            } else {
                level_apply<T>(G, L, x, EpiJacobi<T>{b, x, dinv_of<T>(L), xalt, (T)o.omega});
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
                    // only the very first step of a smoothing call starts from x = 0 (kd == 0)
                    k_cheby0<T><<<grid_for(n, MB, 16384), MB, 0, st>>>(n, dinv_of<T>(L), b, xalt, dv, (T)theta);
                    JFEM_HIP(hipGetLastError());
                } else {
                    EpiCheby<T> e{b, x, dinv_of<T>(L), xalt, dv, (T)theta, (T)c1, (T)c2, kd == 0 ? 1 : 0};
                    level_apply<T>(G, L, x, e);
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
    // with Galerkin coarse systems the boundary conditions are built into them and the
    // transfers apply none (AMGFromMatrix's applyZeroBoundaryConditions is empty)
    const MgLevel& F = G->lev[lf];
    const MgLevel& C = G->lev[lf - 1];
    tp_transfer<T>(G->ctx->stream, F, C, true, xf, xc, G->opt.galerkin || G->opt.raw_transfers ? nullptr : C.bnd, 0);
}

template <typename T>
static void prolong_to(jfem_mg* G, int lf, const T* xc, T* xf, bool add, bool zero_bc)
{
    const MgLevel& F = G->lev[lf];
    const MgLevel& C = G->lev[lf - 1];
    tp_transfer<T>(G->ctx->stream, F, C, false, xc, xf,
                   zero_bc && !G->opt.galerkin && !G->opt.raw_transfers ? F.bnd : nullptr, add ? 1 : 0);
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
    smooth<T>(G, L, b, x, xalt, x_zero, false);
    MgLevel& C = G->lev[l - 1];
    T* r = (T*)L.r;
    level_residual<T>(G, L, b, x, r);
    restrict_to<T>(G, l, r, (T*)C.b);
    step_v<T>(G, l - 1, true);
    prolong_to<T>(G, l, (const T*)C.x, x, true, true);   // guess.add(P e) with zero BC
    if (G->ls.capture && l == (int)G->lev.size() - 1) {
        // training: keep the residual that post-smoothing will see
        level_residual<T>(G, L, b, x, r);
        k_convert<T, double><<<grid_for(L.A->rows, 256, 16384), 256, 0, G->ctx->stream>>>(L.A->rows, r, G->ls.capbuf);
        if (G->ls.capx)
            k_convert<T, double><<<grid_for(L.A->rows, 256, 16384), 256, 0, G->ctx->stream>>>(L.A->rows, x, G->ls.capx);
    }
    smooth<T>(G, L, b, x, xalt, false, true);
    // keep the result in L.x / the spare in L.t
    if (x != (T*)L.x) { L.t = L.x; L.x = x; }
}

// mgStepW (MGPreconditionerInterface.java:200-291): pre-smoothing, then twice the
// coarse correction {defect, restriction, a V-cycle (mgStepV) on level l - 1 from
// zero, prolongation with zero boundary values, addition}, then post-smoothing
template <typename T>
static void step_w(jfem_mg* G, int l, bool x_zero)
{
    MgLevel& L = G->lev[l];
    T* b = (T*)L.b;
    if (l == 0) { coarse_solve<T>(G, b, (T*)L.x); return; }
    T* x = (T*)L.x;
    T* xalt = (T*)L.t;
    smooth<T>(G, L, b, x, xalt, x_zero, false);
    MgLevel& C = G->lev[l - 1];
    T* r = (T*)L.r;
    for (int k = 0; k < 2; k++) {
        level_residual<T>(G, L, b, x, r);
        restrict_to<T>(G, l, r, (T*)C.b);
        step_v<T>(G, l - 1, true);
        prolong_to<T>(G, l, (const T*)C.x, x, true, true);
    }
    smooth<T>(G, L, b, x, xalt, false, true);
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
    if (G->opt.fmg == 2)   // MGPReconditioner2.mvMul: v + fullVCycleSolver(v - A v)
        vec_copy(G->ctx, G->init, rin, n);
    else
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

bool mg_op_apply(const jfem_mat* A, const double* x, double* y)
{
    jfem_mg* G = A->op;
    if (!G || G->lev.back().A != A || !G->lev.back().mf || G->ctx != A->ctx) return false;
    level_apply<double>(G, G->lev.back(), x, EpiStore<double>{y});
    return true;
}

const Grid& mg_finest_grid(const jfem_mg* G) { return G->lev.back().grid; }

const unsigned char* mg_finest_bnd(const jfem_mg* G) { return G->lev.back().bnd; }

hipStream_t mg_stream(const jfem_mg* G) { return G->ctx->stream; }

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
    for (int i = 0; i < its; i++) {
        level_apply<double>(G, L, v, EpiDinv{L.dinv, w});
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
        (void)hipFree(L.tp_buf); (void)hipFree(L.tw0); (void)hipFree(L.tw1);
        (void)hipFree(L.coef); (void)hipFree(L.coef32);
        (void)hipFree(L.mf_cand);
    }
    (void)hipFree(G->coarse_inv);
    (void)hipFree(G->init); (void)hipFree(G->defect); (void)hipFree(G->tmp);
    LearnedState& S = G->ls;
    for (void* p : {(void*)S.stat, (void*)S.stat_dof, (void*)S.slog, (void*)S.ip, (void*)S.perm, (void*)S.rr, (void*)S.u, (void*)S.z, (void*)S.gd, (void*)S.rh, (void*)S.xs,
                    (void*)S.bs, (void*)S.loss, (void*)S.capbuf, (void*)S.capx, (void*)S.xt, (void*)S.eh,
                    (void*)S.w, (void*)S.y, (void*)S.loss2})
        (void)hipFree(p);
    delete G;
}

// ---- Galerkin coarse systems by structured probing ----
// A_c = P^T A P couples coarse nodes I, J only if they share a coarse cell, so
// |I_a - J_a| <= k on every lattice axis.  Colouring the coarse lattice by
// (I_a mod (2k+1)) per axis gives (2k+1)^d colours such that a row's stencil
// holds one node of each: with v_c the indicator of colour c,
// (P^T A P v_c)[I] = A_c[I][J] for the node J of colour c next to I.
__device__ inline int probe_color(const GridDev& g, long dof)
{
    int p[MAXD] = {0, 0, 0};
    lattice_of_dof(g, dof, p);
    const int m = 2 * g.k + 1;
    int c = 0;
    for (int a = g.d - 1; a >= 0; a--) c = c * m + p[a] % m;
    return c;
}

__global__ void k_color_indicator(GridDev g, long n, int color, double* __restrict__ v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        v[i] = probe_color(g, i) == color ? 1.0 : 0.0;
}

// vals[p] of coarse entry (I, J) = Y[colour(J)][I]; entries outside every
// colour's reach stay as they are (none with a structural pattern)
__global__ void k_galerkin_fill(GridDev g, long n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                const double* __restrict__ Y, double* __restrict__ vals)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        for (int p = rowptr[i]; p < rowptr[i + 1]; p++) vals[p] = Y[(long)probe_color(g, cols[p]) * n + i];
}

// A_{lf-1} = P^T A_lf P (FP64, plain P and P^T: no boundary conditions), into
// the coarse level's existing pattern
static void galerkin_level(jfem_mg* G, int lf)
{
    MgLevel& F = G->lev[lf];
    MgLevel& C = G->lev[lf - 1];
    jfem_ctx* c = G->ctx;
    hipStream_t st = c->stream;
    const long nf = F.A->rows, nc = C.A->rows;
    const GridDev gc = to_dev(C.grid);
    int ncol = 1;
    for (int a = 0; a < C.grid.d; a++) ncol *= 2 * C.grid.k + 1;
    DevBuf<double> Y((size_t)nc * ncol), v(nc), w(nf), z(nf);
    for (int col = 0; col < ncol; col++) {
        k_color_indicator<<<grid_for(nc, 256, 16384), 256, 0, st>>>(gc, nc, col, v);
        tp_transfer<double>(st, F, C, false, v, w, nullptr, 0);   // w = P v
        spmv(F.A, w, z);                                          // z = A w
        tp_transfer<double>(st, F, C, true, z, Y.p + (long)col * nc, nullptr, 0);   // P^T z
    }
    k_galerkin_fill<<<grid_for(nc, 256, 16384), 256, 0, st>>>(gc, nc, C.A->rowptr, C.A->colidx, Y, C.A->vals);
    JFEM_HIP(hipGetLastError());
    JFEM_HIP(hipStreamSynchronize(st));
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
            // the transfers' index arithmetic is 32-bit (dof_of_lat, lat_of_dof)
            JFEM_CHECK(L.grid.ndofs() < (1l << 31), JFEM_ERR_UNSUPPORTED, "a multigrid level has 2^31 or more nodes");
            L.A = assemble_matrix(spaces[l], nint, kinds, weights, coeff ? coef[l].data() : nullptr, rule);
            const long n = L.A->rows;
            JFEM_HIP(hipMalloc(&L.bnd, n));
            space_boundary_mask(spaces[l], L.bnd);
            apply_dirichlet_mask(L.A, nullptr, L.bnd, nullptr);
            JFEM_HIP(hipMalloc(&L.dinv, sizeof(double) * n));
            JFEM_HIP(hipMalloc(&L.dinv32, sizeof(float) * n));
            for (void** p : {&L.x, &L.b, &L.r, &L.t, &L.dv}) {
                JFEM_HIP(hipMalloc(p, sizeof(double) * n));
                JFEM_HIP(hipMemsetAsync(*p, 0, sizeof(double) * n, c->stream));
            }
            if (l > 0) build_transfer(L, G->lev[l - 1]);
            // matrix-free level operator: re-assembled systems (every level, or the finest of a
            // Galerkin hierarchy) of GRAD_GRAD / VALUE_VALUE / H1 on 2D / 3D Q1 / Q2 grids
            // matrix_free 2 (auto): 3D grids only, where a CSR row holds 27 (Q1) or 125 (Q2) entries;
            // on 2D grids (9 or 25 per row) the CSR product measured faster
            const bool want = opts->matrix_free == 1 || (opts->matrix_free == 2 && L.grid.d == 3);
            bool mf_ok = want && mf_supported(L.grid) && (!opts->galerkin || l == nlev - 1);
            for (int i = 0; i < nint; i++) mf_ok = mf_ok && kinds[i] >= JFEM_GRAD_GRAD && kinds[i] <= JFEM_H1;
            if (mf_ok) {
                L.mf = true;
                std::memset(&L.mfp, 0, sizeof(L.mfp));
                L.mfp.g = to_dev(L.grid);
                mf_tiles(L.grid, L.mfp.ntile);
                for (int a = 0; a < L.grid.d; a++) element_1d(L.grid, a, rule, L.mfp.K1[a], L.mfp.M1[a]);
                {
                    std::vector<unsigned short> cd;
                    mf_tables(L.grid, cd);
                    if (!cd.empty()) {
                        JFEM_HIP(hipMalloc(&L.mf_cand, sizeof(unsigned short) * cd.size()));
                        JFEM_HIP(hipMemcpy(L.mf_cand, cd.data(), sizeof(unsigned short) * cd.size(),
                                           hipMemcpyHostToDevice));
                        L.mfp.cand = L.mf_cand;
                    }
                }
                for (int i = 0; i < nint; i++) {
                    if (kinds[i] == JFEM_GRAD_GRAD || kinds[i] == JFEM_H1) L.mfp.wgg += weights[i];
                    if (kinds[i] == JFEM_VALUE_VALUE || kinds[i] == JFEM_H1) L.mfp.wvv += weights[i];
                }
                if (coeff) {
                    const long nc = L.grid.ncells();
                    std::vector<float> c32(coef[l].begin(), coef[l].end());
                    JFEM_HIP(hipMalloc(&L.coef, sizeof(double) * nc));
                    JFEM_HIP(hipMalloc(&L.coef32, sizeof(float) * nc));
                    JFEM_HIP(hipMemcpy(L.coef, coef[l].data(), sizeof(double) * nc, hipMemcpyHostToDevice));
                    JFEM_HIP(hipMemcpy(L.coef32, c32.data(), sizeof(float) * nc, hipMemcpyHostToDevice));
                }
            }
        }
        // AMGFromMatrix: coarse systems from the finest one, top down
        if (opts->galerkin)
            for (int l = nlev - 1; l >= 1; l--) galerkin_level(G, l);
        for (int l = 0; l < nlev; l++) {
            MgLevel& L = G->lev[l];
            const long n = L.A->rows;
            if (opts->precision == JFEM_FP32) mat_make_f32(L.A);
            mat_diag(L.A, (double*)L.t);
            k_dinv_from_diag<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(n, (double*)L.t, L.dinv, L.dinv32);
            JFEM_HIP(hipGetLastError());
            if ((opts->smoother == JFEM_SM_CHEBYSHEV || opts->degree > 0) && l > 0)
                L.lmax = (opts->cheb_safety > 0 ? opts->cheb_safety : 1.1) *
                         power_lmax(G, L, opts->power_its > 0 ? opts->power_its : 30);
        }
        // coarsest dense inverse
        MgLevel& C = G->lev[0];
        const int n0 = (int)C.A->rows;
        std::vector<int> rp(n0 + 1), ci(C.A->nnz);
        std::vector<double> cv(C.A->nnz), dense((size_t)n0 * n0, 0.0);
        JFEM_HIP(hipStreamSynchronize(G->ctx->stream));   // blocking copies below do not wait for that stream
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
        if (G->lev.back().mf) G->lev.back().A->op = G;
        JFEM_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
        free_mg(G);
        throw;
    }
    return G;
}


// ------------------------------------------------ learned smoother
// partials of sum (w_i x_i)^2 (w = nullptr: weights 1)
template <typename T>
__global__ __launch_bounds__(256) void k_sumsq(long n, const T* __restrict__ x, const T* __restrict__ w,
                                               double* __restrict__ partials)
{
    __shared__ double sh[4];
    double acc = 0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double v = w ? (double)w[i] * (double)x[i] : (double)x[i];
        acc += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

template <typename T>
static void sumsq_dev(jfem_ctx* c, long n, const T* x, const T* w, double* dst)
{
    const int blocks = grid_for(n, 256, RED_BLOCKS);
    k_sumsq<T><<<blocks, 256, 0, c->stream>>>(n, x, w, c->red_partials);
    JFEM_HIP(hipGetLastError());
    reduce_partials_dev(c, 1, blocks, dst);
}

__device__ inline long pix_of(const GridDev& g, const Lat& L, long i)
{
    int p[MAXD] = {0, 0, 0};
    lattice_of_dof(g, i, p);
    return p[0] + (long)L.n[0] * (p[1] + (long)L.n[1] * p[2]);
}

// lattice index of every DOF (computed once per attached network)
__global__ void k_lattice_perm(GridDev g, Lat L, long n, int* __restrict__ perm)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        perm[i] = (int)pix_of(g, L, i);
}

// network input: channel 0 = s * sqrt(n / ss) with s = D^-1 r the Jacobi-scaled
// residual and ss = |s|^2 (so the input has rms 1 and does not depend on the
// mesh size or the local coefficient scale), channel c >= 1 = channel 0 times
// static channel c (so that the bias-free network is positively homogeneous in r)
template <typename T>
__global__ void k_gather_in(const int* __restrict__ perm, long n, const T* __restrict__ r, const T* __restrict__ dinv,
                            const double* __restrict__ rr, const float* __restrict__ stat, int cin,
                            float* __restrict__ in)
{
    const double scale = rr[0] > 0 ? sqrt((double)n / rr[0]) : 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long pix = perm[i];
        const float v = (float)((double)dinv[i] * (double)r[i] * scale);
        in[pix * cin] = v;
        for (int c = 1; c < cin; c++) in[pix * cin + c] = v * stat[pix * (cin - 1) + c - 1];
    }
}

// the network input of k_gather_in as BF16 for the MFMA first layer: one thread per DOF,
// reading r, D^-1 and the static channels (kept in DOF order for this) contiguously and
// storing the pixel's 16 bytes at its lattice position; the values are those of k_gather_in
// rounded to BF16
template <typename T>
__global__ void k_gather_bf16(const int* __restrict__ perm, long n, const T* __restrict__ r,
                              const T* __restrict__ dinv, const double* __restrict__ rr,
                              const float* __restrict__ stat_dof, int cin, unsigned short* __restrict__ inb)
{
    const double scale = rr[0] > 0 ? sqrt((double)n / rr[0]) : 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float v = (float)((double)dinv[i] * (double)r[i] * scale);
        bf16x8 b;
        b[0] = (short)f2bf(v);
#pragma unroll
        for (int c = 1; c < 8; c++) b[c] = (short)f2bf(c < cin ? v * stat_dof[i * (cin - 1) + c - 1] : 0.f);
        *reinterpret_cast<bf16x8*>(inb + (long)perm[i] * 8) = b;
    }
}

// the network input of one plain learned step, from the residual epilogue of the level
// operator: v = D^-1 (b - A x) in FP32 at the node's lattice position (the first layer forms
// the static channels from it).  The attached network has no biases, so N(c v) = c N(v) for
// c > 0 and the input needs no normalisation: x += N(v) is the step x += rms(v) N(v / rms(v)).
template <typename T>
struct EpiNetIn {
    static constexpr bool reduce = false;
    const T *b, *dinv;
    const int* perm;
    float* vin;
    struct Pre { T b, dinv; int pix; };
    __device__ Pre load(long row) const { return {b[row], dinv[row], perm[row]}; }
    __device__ double operator()(long row, T ax, const Pre& p) const
    {
        vin[p.pix] = (float)(p.dinv * (p.b - ax));
        return 0.0;
    }
};

// the same input for x = 0 (r = b)
template <typename T>
__global__ void k_netin_b(long n, const T* __restrict__ b, const T* __restrict__ dinv, const int* __restrict__ perm,
                          float* __restrict__ vin)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        vin[perm[i]] = (float)(dinv[i] * b[i]);
}

// ip[perm[i]] = i on free DOFs, -1 on Dirichlet DOFs
__global__ void k_pixel_dof(const int* __restrict__ perm, long n, const unsigned char* __restrict__ bnd, int* __restrict__ ip)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        ip[perm[i]] = bnd[i] ? -1 : (int)i;
}

// static channels in DOF order: out[i][c] = stat[perm[i]][c]
__global__ void k_stat_to_dof(long n, int ns, const int* __restrict__ perm, const float* __restrict__ stat,
                              float* __restrict__ out)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        for (int c = 0; c < ns; c++) out[i * ns + c] = stat[(long)perm[i] * ns + c];
}

// x[i] += out[pix] * rms(D^-1 r) on free nodes
template <typename T>
__global__ void k_scatter_add(const int* __restrict__ perm, long n, const float* __restrict__ out, const double* __restrict__ rr,
                              const unsigned char* __restrict__ bnd, T* __restrict__ x)
{
    const double inv = rr[0] > 0 ? sqrt(rr[0] / (double)n) : 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (bnd[i]) continue;
        x[i] = (T)((double)x[i] + (double)out[perm[i]] * inv);
    }
}

// training helpers: u = masked network output (DOF order), rh = r * sqrt(n / ss)
__global__ void k_train_unpack(const int* __restrict__ perm, long n, const float* __restrict__ out, const double* __restrict__ r,
                               const double* __restrict__ rr, const unsigned char* __restrict__ bnd,
                               double* __restrict__ u, double* __restrict__ rh)
{
    const double scale = rr[0] > 0 ? sqrt((double)n / rr[0]) : 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        u[i] = bnd[i] ? 0.0 : (double)out[perm[i]];
        rh[i] = r[i] * scale;
    }
}

// dLoss/dout for the energy loss (u.A u - 2 u.rh) / n: gout = -2 z / n on free nodes, z = rh - A u
__global__ void k_train_gout(const int* __restrict__ perm, long n, const double* __restrict__ z,
                             const unsigned char* __restrict__ bnd, float* __restrict__ gout)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        gout[perm[i]] = bnd[i] ? 0.f : (float)(-2.0 * z[i] / (double)n);
}

// partials of -u.(z + rh)
__global__ __launch_bounds__(256) void k_energy(long n, const double* __restrict__ u, const double* __restrict__ z,
                                                const double* __restrict__ rh, double* __restrict__ partials)
{
    __shared__ double sh[4];
    double acc = 0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) acc -= u[i] * (z[i] + rh[i]);
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// eh = (xt - x) * sqrt(n / ss): the error of the sample in the input's units
__global__ void k_scaled_error(long n, const double* __restrict__ xt, const double* __restrict__ x,
                               const double* __restrict__ rr, double* __restrict__ eh)
{
    const double scale = rr[0] > 0 ? sqrt((double)n / rr[0]) : 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        eh[i] = (xt[i] - x[i]) * scale;
}

// partials of (eh - u - w).y (energy of the two-grid error) and eh.rh (energy of the error)
__global__ __launch_bounds__(256) void k_twogrid_dots(long n, const double* __restrict__ eh, const double* __restrict__ u,
                                                      const double* __restrict__ w, const double* __restrict__ y,
                                                      const double* __restrict__ rh, double* __restrict__ partials)
{
    __shared__ double sh[2][4];
    double a = 0, b = 0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        a += (eh[i] - u[i] - w[i]) * y[i];
        b += eh[i] * rh[i];
    }
    for (int o = 32; o > 0; o >>= 1) { a += __shfl_xor(a, o, 64); b += __shfl_xor(b, o, 64); }
    if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = a; sh[1][threadIdx.x >> 6] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
        partials[RED_BLOCKS + blockIdx.x] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    }
}

// seeded standard-normal right-hand side (SplitMix64 + Box-Muller), zero on the boundary
__global__ void k_rand_rhs(long n, unsigned long long seed, const unsigned char* __restrict__ bnd, double* __restrict__ b)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        auto mix = [](unsigned long long z) {
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        const unsigned long long z1 = mix(seed + (2 * (unsigned long long)i + 1) * 0x9E3779B97F4A7C15ull);
        const unsigned long long z2 = mix(seed + (2 * (unsigned long long)i + 2) * 0x9E3779B97F4A7C15ull);
        const double u1 = ((z1 >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        const double u2 = (z2 >> 11) * (1.0 / 9007199254740992.0);
        b[i] = bnd[i] ? 0.0 : sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
}

// u = network output on free nodes (DOF order)
template <typename T>
__global__ void k_unpack_out(const int* __restrict__ perm, long n, const float* __restrict__ out, const unsigned char* __restrict__ bnd,
                             T* __restrict__ u)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        u[i] = bnd[i] ? (T)0 : (T)out[perm[i]];
}

// partials of (r.w), (w.w) (minimal residual, energy = 0) or (u.r), (u.w)
// (minimal energy, energy = 1) into slots 0 and 1
template <typename T>
__global__ __launch_bounds__(256) void k_dot2(long n, const T* __restrict__ r, const T* __restrict__ w,
                                              const T* __restrict__ u, int energy, double* __restrict__ partials)
{
    __shared__ double sh[2][4];
    double a = 0, c = 0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double wi = (double)w[i];
        if (energy) {
            const double ui = (double)u[i];
            a += ui * (double)r[i];
            c += ui * wi;
        } else {
            a += (double)r[i] * wi;
            c += wi * wi;
        }
    }
    for (int o = 32; o > 0; o >>= 1) { a += __shfl_xor(a, o, 64); c += __shfl_xor(c, o, 64); }
    if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = a; sh[1][threadIdx.x >> 6] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
        partials[RED_BLOCKS + blockIdx.x] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    }
}

// x += alpha u with alpha = (r.Au)/(Au.Au) (the step never increases ||r||_2) or
// alpha = (u.r)/(u.Au) (it never increases the energy norm of the error)
template <typename T>
__global__ void k_mr_update(long n, const double* __restrict__ d, const T* __restrict__ u, T* __restrict__ x)
{
    const double alpha = d[1] > 0 ? d[0] / d[1] : 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = (T)((double)x[i] + alpha * (double)u[i]);
}

// learned smoothing step: r = b - A x, u = N(r / rms r) (direction), then a
// minimal-residual step along u
template <typename T>
static void learned_smooth(jfem_mg* G, MgLevel& L, const T* b, T* x, T* u, bool x_zero)
{
    LearnedState& S = G->ls;
    jfem_ctx* c = G->ctx;
    hipStream_t st = c->stream;
    const long n = L.A->rows;
    const int gr = grid_for(n, 256, 16384);
    T* r = (T*)L.r;
    T* w = (T*)L.dv;
    for (int it = 0; it < S.runs; it++) {
        if (!S.mr && nn_bf16_input(S.nn, S.lat)) {
            // plain step on the MFMA network: the residual epilogue writes v at the nodes'
            // lattice positions; the last layer adds N(v) into x on the free nodes
            if (x_zero && it == 0) {
                k_fill_t<T><<<gr, 256, 0, st>>>(n, (T)0, x);
                k_netin_b<T><<<gr, 256, 0, st>>>(n, b, dinv_of<T>(L), S.perm, S.nn->in);
            } else {
                level_apply<T>(G, L, x, EpiNetIn<T>{b, dinv_of<T>(L), S.perm, S.nn->in});
            }
            JFEM_HIP(hipGetLastError());
            nn_forward_plain(S.nn, S.lat, S.slog, S.ip, x, sizeof(T) == 8);
            continue;
        }
        if (x_zero && it == 0) {
            k_fill_t<T><<<gr, 256, 0, st>>>(n, (T)0, x);
            JFEM_HIP(hipMemcpyAsync(r, b, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
        } else {
            level_residual<T>(G, L, b, x, r);
        }
        sumsq_dev<T>(c, n, r, dinv_of<T>(L), S.rr);
        if (nn_bf16_input(S.nn, S.lat)) {
            k_gather_bf16<T><<<gr, 256, 0, st>>>(S.perm, n, r, dinv_of<T>(L), S.rr, S.stat_dof, S.nn->cin,
                                                 S.nn->inb);
            nn_forward(S.nn, S.lat, true);
        } else {
            k_gather_in<T><<<gr, 256, 0, st>>>(S.perm, n, r, dinv_of<T>(L), S.rr, S.stat, S.nn->cin,
                                               S.nn->in);
            nn_forward(S.nn, S.lat);
        }
        if (!S.mr) {
            k_scatter_add<T><<<gr, 256, 0, st>>>(S.perm, n, S.nn->out, S.rr, L.bnd, x);
            JFEM_HIP(hipGetLastError());
            continue;
        }
        k_unpack_out<T><<<gr, 256, 0, st>>>(S.perm, n, S.nn->out, L.bnd, u);
        level_apply<T>(G, L, u, EpiStore<T>{w});
        const int blocks = grid_for(n, 256, RED_BLOCKS);
        k_dot2<T><<<blocks, 256, 0, st>>>(n, r, w, u, S.mr == 2 ? 1 : 0, c->red_partials);
        reduce_partials_dev(c, 2, blocks, S.rr + 2);
        k_mr_update<T><<<gr, 256, 0, st>>>(n, S.rr + 2, u, x);
        JFEM_HIP(hipGetLastError());
    }
}

// w = P V_c(P^T z): the coarse-grid correction of the finest level for the
// residual z (restriction and prolongation with zero boundary conditions, one
// V-cycle from zero on the coarser levels), in the hierarchy's precision
template <typename T>
static void coarse_correction_t(jfem_mg* G, const double* z, double* w)
{
    const int Lf = (int)G->lev.size() - 1;
    MgLevel& F = G->lev[Lf];
    MgLevel& C = G->lev[Lf - 1];
    const long n = F.A->rows;
    hipStream_t st = G->ctx->stream;
    const int g = grid_for(n, 256, 16384);
    T* zt = (T*)F.r;
    T* wt = (T*)F.dv;
    k_convert<double, T><<<g, 256, 0, st>>>(n, z, zt);
    restrict_to<T>(G, Lf, zt, (T*)C.b);
    step_v<T>(G, Lf - 1, true);
    prolong_to<T>(G, Lf, (const T*)C.x, wt, false, true);
    k_convert<T, double><<<g, 256, 0, st>>>(n, wt, w);
    JFEM_HIP(hipGetLastError());
}

static void coarse_correction(jfem_mg* G, const double* z, double* w)
{
    if (G->opt.precision == JFEM_FP32) coarse_correction_t<float>(G, z, w);
    else coarse_correction_t<double>(G, z, w);
}

// one V-cycle x <- mgStepV(finest, x, b) on FP64 device vectors, any precision inside
static void vcycle_dev(jfem_mg* G, double* x, const double* b)
{
    const int Lf = (int)G->lev.size() - 1;
    MgLevel& F = G->lev[Lf];
    const long n = F.A->rows;
    hipStream_t st = G->ctx->stream;
    const int g = grid_for(n, 256, 16384);
    if (G->opt.precision == JFEM_FP32) {
        k_convert<double, float><<<g, 256, 0, st>>>(n, b, (float*)F.b);
        k_convert<double, float><<<g, 256, 0, st>>>(n, x, (float*)F.x);
        step_v<float>(G, Lf, false);
        k_convert<float, double><<<g, 256, 0, st>>>(n, (const float*)F.x, x);
    } else {
        vec_copy(G->ctx, (double*)F.b, b, n);
        vec_copy(G->ctx, (double*)F.x, x, n);
        step_v<double>(G, Lf, false);
        vec_copy(G->ctx, x, (const double*)F.x, n);
    }
    JFEM_HIP(hipGetLastError());
}

void mg_observe_reduction(jfem_mg* G, double ratio)
{
    if (G->ls.active && ratio > G->ls.fallback) {
        G->ls.active = 0;
        G->ls.fell_back = 1;
    }
}

static void set_learned(jfem_mg* G, jfem_nn* nn, const double* cell_logc, int runs, double fallback, int mr)
{
    LearnedState& S = G->ls;
    if (!nn) {
        S.nn = nullptr;
        S.active = 0;
        return;
    }
    MgLevel& F = G->lev.back();
    const Grid& g = F.grid;
    JFEM_CHECK(nn->dim == g.d, JFEM_ERR_ARG, "network dimension differs from the grid");
    JFEM_CHECK(nn->cin == 2 || nn->cin == 2 + g.d, JFEM_ERR_ARG, "network needs 2 or 2+dim input channels");
    JFEM_CHECK(F.lmax > 0, JFEM_ERR_STATE, "hierarchy has no Chebyshev fallback (create it with a Chebyshev smoother)");
    int dims[3] = {1, 1, 1};
    for (int a = 0; a < g.d; a++) dims[a] = (int)g.nodes_axis(a);
    S.lat = make_lat(g.d, dims);
    const long np = S.lat.np;
    JFEM_CHECK(np == F.A->rows, JFEM_ERR_STATE, "lattice and DOF counts differ");
    const int ns = nn->cin - 1;
    std::vector<float> stat((size_t)np * ns, 0.f);
    // channel 1: mean log-coefficient of the cells around each node
    if (cell_logc) {
        for (long pix = 0; pix < np; pix++) {
            int p[3] = {(int)(pix % dims[0]), (int)((pix / dims[0]) % dims[1]), (int)(pix / ((long)dims[0] * dims[1]))};
            int lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                if (a >= g.d) { lo[a] = hi[a] = 0; continue; }
                if (p[a] % g.k == 0) { lo[a] = p[a] / g.k - 1; hi[a] = p[a] / g.k; } else { lo[a] = hi[a] = p[a] / g.k; }
                lo[a] = std::max(lo[a], 0);
                hi[a] = std::min(hi[a], g.n[a] - 1);
            }
            double s = 0;
            int cnt = 0;
            for (int c2 = lo[2]; c2 <= hi[2]; c2++)
                for (int c1 = lo[1]; c1 <= hi[1]; c1++)
                    for (int c0 = lo[0]; c0 <= hi[0]; c0++) {
                        const int cc[3] = {c0, c1, c2};
                        long cell = 0;
                        for (int a = 0; a < g.d; a++) cell = cell * g.n[a] + cc[a];
                        s += cell_logc[cell];
                        cnt++;
                    }
            stat[(size_t)pix * ns] = (float)(s / cnt);
        }
    }
    // channels 2..: lattice parity per axis (node position inside its cell for Q2)
    if (nn->cin == 2 + g.d)
        for (long pix = 0; pix < np; pix++) {
            long r = pix;
            for (int a = 0; a < g.d; a++) {
                stat[(size_t)pix * ns + 1 + a] = (float)((r % dims[a]) % 2);
                r /= dims[a];
            }
        }
    (void)hipFree(S.stat);
    S.stat = nullptr;
    JFEM_HIP(hipMalloc(&S.stat, sizeof(float) * stat.size()));
    JFEM_HIP(hipMemcpy(S.stat, stat.data(), sizeof(float) * stat.size(), hipMemcpyHostToDevice));
    (void)hipFree(S.perm);
    S.perm = nullptr;
    JFEM_HIP(hipMalloc(&S.perm, sizeof(int) * np));
    k_lattice_perm<<<grid_for(np, 256, 16384), 256, 0, G->ctx->stream>>>(to_dev(g), S.lat, np, S.perm);
    JFEM_HIP(hipGetLastError());
    (void)hipFree(S.stat_dof);
    S.stat_dof = nullptr;
    JFEM_HIP(hipMalloc(&S.stat_dof, sizeof(float) * stat.size()));
    k_stat_to_dof<<<grid_for(np, 256, 16384), 256, 0, G->ctx->stream>>>(np, ns, S.perm, S.stat, S.stat_dof);
    JFEM_HIP(hipGetLastError());
    (void)hipFree(S.ip);
    S.ip = nullptr;
    JFEM_HIP(hipMalloc(&S.ip, sizeof(int) * np));
    k_pixel_dof<<<grid_for(np, 256, 16384), 256, 0, G->ctx->stream>>>(S.perm, np, F.bnd, S.ip);
    JFEM_HIP(hipGetLastError());
    (void)hipFree(S.slog);
    S.slog = nullptr;
    if (cell_logc) {
        std::vector<float> sl((size_t)np);
        for (long pix = 0; pix < np; pix++) sl[(size_t)pix] = stat[(size_t)pix * ns];
        JFEM_HIP(hipMalloc(&S.slog, sizeof(float) * np));
        JFEM_HIP(hipMemcpy(S.slog, sl.data(), sizeof(float) * np, hipMemcpyHostToDevice));
    }
    if (!S.rr) JFEM_HIP(hipMalloc(&S.rr, sizeof(double) * 8));
    nn_reserve(nn, np);
    nn_set_bias_free(nn, 1);   // N(0) = 0 and N(c r) = c N(r): a consistent smoother
    S.nn = nn;
    S.runs = runs > 0 ? runs : 1;
    S.mr = (mr >= 0 && mr <= 2) ? mr : 0;
    S.fallback = fallback > 0 ? fallback : 0.5;
    S.active = 1;
    S.fell_back = 0;
}

// Train the attached network on residuals of plain V-cycles of this hierarchy:
// for each seeded random right-hand side b (white noise, or for odd b smooth:
// noise three levels down, prolongated), x = 0, and for each cycle k the
// sample is r_k = b - A x_k with error e_k = A^-1 r_k; with the input
// normalised to rms 1 the loss is the energy of the error left by the step,
// |e_k - N(r_k)|_A^2 - |e_k|_A^2 = u.A u - 2 u.r (per node), which needs only
// r_k; then x_{k+1} = V(x_k) with the plain smoother.
static int train_mg(jfem_mg* G, const jfem_train_opts* o, double* loss_host)
{
    LearnedState& S = G->ls;
    JFEM_CHECK(S.nn, JFEM_ERR_STATE, "no network attached (jfem_mg_set_learned)");
    jfem_ctx* c = G->ctx;
    hipStream_t st = c->stream;
    MgLevel& F = G->lev.back();
    const long n = F.A->rows;
    const int gr = grid_for(n, 256, 16384);
    const int nsteps = o->rhs_per_field * o->cycles_per_rhs;
    const bool twogrid = o->mode == 4 || o->mode == 5;
    const bool onpolicy = o->mode == 5;   // odd right-hand sides advance with the learned V-cycle
    JFEM_CHECK(!twogrid || G->lev.size() >= 2, JFEM_ERR_STATE, "two-grid training needs at least two levels");
    for (double** p : {&S.u, &S.z, &S.gd, &S.rh, &S.xs, &S.bs, &S.capbuf}) {
        (void)hipFree(*p);
        *p = nullptr;
        JFEM_HIP(hipMalloc(p, sizeof(double) * n));
    }
    for (double** p : {&S.capx, &S.xt, &S.eh, &S.w, &S.y}) {
        (void)hipFree(*p);
        *p = nullptr;
        if (twogrid) JFEM_HIP(hipMalloc(p, sizeof(double) * n));
    }
    if (nsteps > S.loss_cap) {
        (void)hipFree(S.loss);
        JFEM_HIP(hipMalloc(&S.loss, sizeof(double) * nsteps));
        S.loss_cap = nsteps;
    }
    if (twogrid && 2L * nsteps > S.loss2_cap) {
        (void)hipFree(S.loss2);
        JFEM_HIP(hipMalloc(&S.loss2, sizeof(double) * 2 * nsteps));
        S.loss2_cap = 2L * nsteps;
    }
    const int was_active = S.active;
    S.active = 0;   // data come from plain V-cycles
    double* r = G->tmp;
    int step = 0;
    const int Lf = (int)G->lev.size() - 1;
    for (int k = 0; k < o->rhs_per_field; k++) {
        const unsigned long long sk = o->seed * 0x100000001B3ull + (unsigned long long)k * 7919ull;
        if ((k & 1) && Lf >= 1) {
            // smooth right-hand side: white noise three levels down, prolongated to the finest grid
            const int ls = Lf >= 3 ? Lf - 3 : 0;
            MgLevel& C = G->lev[ls];
            k_rand_rhs<<<grid_for(C.A->rows, 256, 16384), 256, 0, st>>>(C.A->rows, sk, C.bnd, (double*)C.t);
            for (int l = ls + 1; l <= Lf; l++)
                prolong_to<double>(G, l, (const double*)G->lev[l - 1].t, (double*)G->lev[l].t, false, true);
            vec_copy(c, twogrid ? S.xt : S.bs, (const double*)F.t, n);
        } else {
            k_rand_rhs<<<gr, 256, 0, st>>>(n, sk, F.bnd, twogrid ? S.xt : S.bs);
        }
        // two-grid training: the random vector is the exact solution x*, b = A x*
        if (twogrid) spmv(F.A, S.xt, S.bs);
        vec_fill(c, S.xs, 0.0, n);
        for (int cyc = 0; cyc < o->cycles_per_rhs; cyc++) {
            const bool noise = o->mode == 1 || (o->mode == 2 && (step & 1));
            const bool captured = (o->mode == 3 || twogrid) && (step & 1);
            if (captured) {
                vec_copy(c, r, S.capbuf, n);
            } else if (noise) {
                // r = A e for a white-noise error e (zero on the boundary)
                k_rand_rhs<<<gr, 256, 0, st>>>(n, o->seed * 0x9E3779B97F4A7C15ull + 0x51ull * (step + 1), F.bnd, S.gd);
                spmv(F.A, S.gd, r);
            } else {
                level_apply<double>(G, F, S.xs, EpiResidual<double>{S.bs, r});
            }
            sumsq_dev<double>(c, n, r, F.dinv, S.rr);
            k_gather_in<double><<<gr, 256, 0, st>>>(S.perm, n, r, F.dinv, S.rr, S.stat, S.nn->cin,
                                                    S.nn->in);
            nn_forward(S.nn, S.lat);
            k_train_unpack<<<gr, 256, 0, st>>>(S.perm, n, S.nn->out, r, S.rr, F.bnd, S.u, S.rh);
            level_apply<double>(G, F, S.u, EpiResidual<double>{S.rh, S.z});   // z = rh - A u
            const int blocks = grid_for(n, 256, RED_BLOCKS);
            if (twogrid) {
                // two-grid loss: energy of e2 = T (e - u), T = I - P A_c^-1 P^T A the coarse-grid
                // correction (A_c^-1 approximated by one V-cycle); with T an A-orthogonal
                // projector the gradient is -2 A e2 = -2 y, y = z - A w, w = P A_c^-1 P^T z
                k_scaled_error<<<gr, 256, 0, st>>>(n, S.xt, captured ? S.capx : S.xs, S.rr, S.eh);
                coarse_correction(G, S.z, S.w);
                level_apply<double>(G, F, S.w, EpiResidual<double>{S.z, S.y});
                k_twogrid_dots<<<blocks, 256, 0, st>>>(n, S.eh, S.u, S.w, S.y, S.rh, c->red_partials);
                reduce_partials_dev(c, 2, blocks, S.loss2 + 2 * step);
                k_train_gout<<<gr, 256, 0, st>>>(S.perm, n, S.y, F.bnd, S.nn->gout);
            } else {
                // energy loss |e - u|_A^2 - |e|_A^2 = u.A u - 2 u.rh = -u.(z + rh), gradient 2 (A u - rh) = -2 z
                k_energy<<<blocks, 256, 0, st>>>(n, S.u, S.z, S.rh, c->red_partials);
                reduce_partials_dev(c, 1, blocks, S.loss + step);
                k_train_gout<<<gr, 256, 0, st>>>(S.perm, n, S.z, F.bnd, S.nn->gout);
            }
            JFEM_HIP(hipGetLastError());
            nn_backward(S.nn, S.lat, 0);
            nn_adam_step(S.nn, (float)o->lr, 1.0f);
            step++;
            if (o->mode == 3 || twogrid) {
                // even steps train on the cycle-entry residual; the V-cycle that
                // follows records the residual before post-smoothing for the odd step
                if (!captured) {
                    S.capture = 1;
                    S.active = onpolicy && (k & 1) ? 1 : 0;
                    vcycle_dev(G, S.xs, S.bs);
                    S.active = 0;
                    S.capture = 0;
                }
            } else if (!noise && cyc + 1 < o->cycles_per_rhs) {
                vcycle_dev(G, S.xs, S.bs);
            }
        }
    }
    if (twogrid) {
        // per step: squared two-grid energy contraction of the sample, |e2|_A^2 / |e|_A^2
        std::vector<double> h(2 * (size_t)step);
        JFEM_HIP(hipMemcpyAsync(h.data(), S.loss2, sizeof(double) * 2 * step, hipMemcpyDeviceToHost, st));
        JFEM_HIP(hipStreamSynchronize(st));
        if (loss_host)
            for (int i = 0; i < step; i++) loss_host[i] = h[2 * i + 1] > 0 ? h[2 * i] / h[2 * i + 1] : 0.0;
    } else {
        std::vector<double> h(step);
        JFEM_HIP(hipMemcpyAsync(h.data(), S.loss, sizeof(double) * step, hipMemcpyDeviceToHost, st));
        JFEM_HIP(hipStreamSynchronize(st));
        if (loss_host)
            for (int i = 0; i < step; i++) loss_host[i] = h[i] / (double)n;
    }
    S.active = was_active;
    return step;
}

// stationary multigrid iteration x <- V(x) until ||b - A x|| <= tol (FP64
// residual on the finest matrix), with the learned smoother's fallback rule
static void mg_solve(jfem_mg* G, const double* b, double* x, double tol, int maxcycles, jfem_solve_info* info,
                     double* hist)
{
    MgLevel& F = G->lev.back();
    jfem_ctx* c = G->ctx;
    const long n = F.A->rows;
    double* r = G->tmp;
    launch_stream<double, double>(F.A, F.A->vals, x, EpiResidual<double>{b, r}, nullptr);
    double rn = std::sqrt(vec_dot(c, n, r, r));
    int it = 0;
    while (rn > tol && it < maxcycles) {
        vcycle_dev(G, x, b);
        launch_stream<double, double>(F.A, F.A->vals, x, EpiResidual<double>{b, r}, nullptr);
        const double rnew = std::sqrt(vec_dot(c, n, r, r));
        JFEM_CHECK(std::isfinite(rnew), JFEM_ERR_STATE, "multigrid iteration diverged");
        mg_observe_reduction(G, rn > 0 ? rnew / rn : 0.0);
        rn = rnew;
        if (hist) hist[it] = rn;
        it++;
    }
    info->iterations = it;
    info->resnorm = rn;
    info->fell_back = G->ls.fell_back;
}

}  // namespace jfem

using namespace jfem;

extern "C" {

int jfem_mg_create(jfem_space* const* spaces, int nlev, int nint, const int* kinds, const double* weights,
                   const double* cell_coeff, int quad_rule, const jfem_mg_opts* opts, jfem_mg** out)
{
    API_BEGIN
    JFEM_CHECK(out && kinds && weights, JFEM_ERR_ARG, "null argument");
    for (int i = 0; i < nint; i++)
        JFEM_CHECK(kinds[i] >= 0 && kinds[i] <= 2, JFEM_ERR_UNSUPPORTED,
                   "multigrid hierarchies take GRAD_GRAD, VALUE_VALUE and H1 integrals");
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

int jfem_mg_level_apply(jfem_mg* mg, int level, const jfem_vec* x, jfem_vec* y)
{
    API_BEGIN
    JFEM_CHECK(mg && x && y, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(level >= 0 && level < (int)mg->lev.size(), JFEM_ERR_ARG, "no such level");
    const MgLevel& L = mg->lev[level];
    JFEM_CHECK(x->n == L.A->rows && y->n == L.A->rows, JFEM_ERR_ARG, "Incompatible sizes");
    JFEM_CHECK(x->d != y->d || L.A->rows == 0, JFEM_ERR_ARG, "output aliases the input");
    level_apply<double>(mg, L, x->d, EpiStore<double>{y->d});
    JFEM_HIP(hipStreamSynchronize(mg->ctx->stream));
    API_END
}

int jfem_mg_level_bench(jfem_mg* mg, int level, const jfem_vec* x, jfem_vec* y, int reps, double* ms)
{
    API_BEGIN
    JFEM_CHECK(mg && x && y && ms && reps > 0, JFEM_ERR_ARG, "bad argument");
    JFEM_CHECK(level >= 0 && level < (int)mg->lev.size(), JFEM_ERR_ARG, "no such level");
    MgLevel& L = mg->lev[level];
    JFEM_CHECK(x->n == L.A->rows && y->n == L.A->rows, JFEM_ERR_ARG, "Incompatible sizes");
    JFEM_CHECK(x->d != y->d || L.A->rows == 0, JFEM_ERR_ARG, "output aliases the input");
    JFEM_CHECK(mg->opt.precision == JFEM_FP64, JFEM_ERR_UNSUPPORTED, "the step is timed on FP64 hierarchies");
    // a non-first Chebyshev step (d = c1 d + c2 D^-1 (b - A x), y = x + d) with the level's own
    // right-hand side and direction work vectors as b and d
    const EpiCheby<double> e{(const double*)L.b, x->d, L.dinv, y->d, (double*)L.dv, 1.0, 0.5, 0.5, 0};
    hipStream_t st = mg->ctx->stream;
    hipEvent_t e0, e1;
    JFEM_HIP(hipEventCreate(&e0));
    JFEM_HIP(hipEventCreate(&e1));
    level_apply<double>(mg, L, x->d, e);   // warm-up launch
    JFEM_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) level_apply<double>(mg, L, x->d, e);
    JFEM_HIP(hipEventRecord(e1, st));
    JFEM_HIP(hipEventSynchronize(e1));
    float f = 0;
    JFEM_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    API_END
}

int jfem_mg_level_matrix_free(jfem_mg* mg, int level, int* out)
{
    API_BEGIN
    JFEM_CHECK(mg && out, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(level >= 0 && level < (int)mg->lev.size(), JFEM_ERR_ARG, "no such level");
    *out = mg->lev[level].mf ? 1 : 0;
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

int jfem_mg_prolongate(jfem_mg* mg, int level, const jfem_vec* coarse, jfem_vec* fine)
{
    API_BEGIN
    JFEM_CHECK(mg && coarse && fine, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(level >= 0 && level + 1 < (int)mg->lev.size(), JFEM_ERR_ARG, "level out of range");
    const MgLevel& F = mg->lev[level + 1];
    const MgLevel& C = mg->lev[level];
    JFEM_CHECK(coarse->n == C.A->rows && fine->n == F.A->rows, JFEM_ERR_ARG, "wrong size");
    tp_transfer<double>(mg->ctx->stream, F, C, false, coarse->d, fine->d, nullptr, 0);
    API_END
}

int jfem_mg_restrict(jfem_mg* mg, int level, const jfem_vec* fine, jfem_vec* coarse)
{
    API_BEGIN
    JFEM_CHECK(mg && coarse && fine, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(level >= 0 && level + 1 < (int)mg->lev.size(), JFEM_ERR_ARG, "level out of range");
    const MgLevel& F = mg->lev[level + 1];
    const MgLevel& C = mg->lev[level];
    JFEM_CHECK(coarse->n == C.A->rows && fine->n == F.A->rows, JFEM_ERR_ARG, "wrong size");
    tp_transfer<double>(mg->ctx->stream, F, C, true, fine->d, coarse->d, nullptr, 0);
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

int jfem_mg_wcycle(jfem_mg* mg, jfem_vec* x, const jfem_vec* b)
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
        step_w<float>(mg, Lf, false);
        k_convert<float, double><<<g, 256, 0, st>>>(n, (const float*)F.x, x->d);
    } else {
        vec_copy(mg->ctx, (double*)F.b, b->d, n);
        vec_copy(mg->ctx, (double*)F.x, x->d, n);
        step_w<double>(mg, Lf, false);
        vec_copy(mg->ctx, x->d, (const double*)F.x, n);
    }
    JFEM_HIP(hipGetLastError());
    JFEM_HIP(hipStreamSynchronize(st));
    API_END
}

int jfem_mg_set_learned(jfem_mg* mg, jfem_nn* nn, const double* cell_log_coeff, int runs, double fallback_ratio,
                        int step_rule)
{
    API_BEGIN
    JFEM_CHECK(mg, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(step_rule >= 0 && step_rule <= 2, JFEM_ERR_ARG, "unknown step rule");
    set_learned(mg, nn, cell_log_coeff, runs, fallback_ratio, step_rule);
    JFEM_HIP(hipStreamSynchronize(mg->ctx->stream));
    API_END
}

int jfem_mg_learned_status(jfem_mg* mg, int* active, int* fell_back)
{
    API_BEGIN
    JFEM_CHECK(mg, JFEM_ERR_ARG, "null argument");
    if (active) *active = mg->ls.active;
    if (fell_back) *fell_back = mg->ls.fell_back;
    API_END
}

int jfem_mg_set_learned_sides(jfem_mg* mg, int sides)
{
    API_BEGIN
    JFEM_CHECK(mg, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(sides >= 1 && sides <= 3, JFEM_ERR_ARG, "sides must be 1 (pre), 2 (post) or 3 (both)");
    mg->ls.sides = sides;
    API_END
}

int jfem_mg_set_learned_active(jfem_mg* mg, int active)
{
    API_BEGIN
    JFEM_CHECK(mg, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(!active || mg->ls.nn, JFEM_ERR_STATE, "no network attached");
    mg->ls.active = active ? 1 : 0;
    mg->ls.fell_back = 0;
    API_END
}

int jfem_nn_train_mg(jfem_mg* mg, const jfem_train_opts* opts, double* loss_hist, int* nsteps)
{
    API_BEGIN
    JFEM_CHECK(mg && opts, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(opts->rhs_per_field >= 1 && opts->cycles_per_rhs >= 1 && opts->lr > 0, JFEM_ERR_ARG,
               "bad training options");
    const int s = train_mg(mg, opts, loss_hist);
    if (nsteps) *nsteps = s;
    API_END
}

int jfem_mg_solve(jfem_mg* mg, const jfem_vec* b, jfem_vec* x, double tol, int maxcycles, jfem_solve_info* info,
                  double* res_hist)
{
    API_BEGIN
    JFEM_CHECK(mg && b && x && info, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(b->n == mg_size(mg) && x->n == mg_size(mg), JFEM_ERR_ARG, "wrong size");
    std::memset(info, 0, sizeof(*info));
    JFEM_HIP(hipStreamSynchronize(mg->ctx->stream));
    auto t0 = std::chrono::steady_clock::now();
    mg_solve(mg, b->d, x->d, tol, maxcycles, info, res_hist);
    JFEM_HIP(hipStreamSynchronize(mg->ctx->stream));
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
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
