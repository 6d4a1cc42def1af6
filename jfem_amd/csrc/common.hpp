// Shared definitions of libjfem_hip: context, handles, error reporting,
// launch helpers and the structured-grid DOF arithmetic used by every kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/jfem_hip.h"

namespace jfem {

// ---------------------------------------------------------------------------
// errors: every C-ABI call returns a jfem_status; the message of the last
// failure on the calling thread is kept for jfem_last_error().
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define JFEM_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::jfem::Error{JFEM_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define API_BEGIN try {
#define API_END                                                   \
    }                                                             \
    catch (const ::jfem::Error& e) {                              \
        ::jfem::set_error(e.msg);                                 \
        return e.code;                                            \
    }                                                             \
    catch (const std::bad_alloc&) {                               \
        ::jfem::set_error("host allocation failed");              \
        return JFEM_ERR_NOMEM;                                    \
    }                                                             \
    catch (...) {                                                 \
        ::jfem::set_error("unknown failure");                     \
        return JFEM_ERR_STATE;                                    \
    }                                                             \
    return JFEM_OK;

#define JFEM_CHECK(cond, code, msg)                                 \
    do {                                                            \
        if (!(cond)) throw ::jfem::Error{(code), std::string(msg)}; \
    } while (0)

// ---------------------------------------------------------------------------
// structured Qk space on a Cartesian box (ContinuousTPFESpace)
// ---------------------------------------------------------------------------
constexpr int MAXD = 3;

struct Grid {
    int d = 0, k = 0;
    int n[MAXD] = {1, 1, 1};   // cells per axis
    double start[MAXD] = {0, 0, 0}, end[MAXD] = {0, 0, 0};
    long ncells() const { long c = 1; for (int a = 0; a < d; a++) c *= n[a]; return c; }
    long nodes_axis(int a) const { return (long)k * n[a] + 1; }
    long ndofs() const { long c = 1; for (int a = 0; a < d; a++) c *= nodes_axis(a); return c; }
    int nloc() const { int c = 1; for (int a = 0; a < d; a++) c *= k + 1; return c; }
};

// Closed form of the reference's first-appearance DOF numbering
// (ContinuousTPFESpace.assembleFunctions, ContinuousTPFESpace.java:29-50):
// cells are visited in IntCoordinates order (last axis fastest), the local
// nodes of a cell with axis 0 fastest, and a node is new in cell c iff l_a >= 1
// on every axis with c_a > 0.  Hence with m_a(c) = (c > 0 ? k : k + 1),
// S_a(c) = (c > 0 ? c*k + 1 : 0) and T_a = k*n_a + 1:
//   prefix(c) = sum_a [prod_{b<a} m_b(c_b)] * S_a(c_a) * [prod_{b>a} T_b]
//   rank(l|c) = sum_a (l_a - [c_a > 0]) * prod_{b<a} m_b(c_b)
// and the node at lattice point p is owned by c_a = (p_a == 0 ? 0 : (p_a-1)/k).
struct GridDev {
    int d, k;
    int n[MAXD];
};

__host__ __device__ inline long dof_of_lattice(const GridDev& g, const int* p)
{
    long pre = 0, mprod = 1, rank = 0;
    // tail products T_{b>a}
    long tail[MAXD];
    long t = 1;
    for (int a = g.d - 1; a >= 0; a--) { tail[a] = t; t *= (long)g.k * g.n[a] + 1; }
    for (int a = 0; a < g.d; a++) {
        const int c = p[a] == 0 ? 0 : (p[a] - 1) / g.k;
        const int l = p[a] - c * g.k;
        const long m = c > 0 ? g.k : g.k + 1;
        const long S = c > 0 ? (long)c * g.k + 1 : 0;
        pre += mprod * S * tail[a];
        rank += (long)(l - (c > 0 ? 1 : 0)) * mprod;
        mprod *= m;
    }
    return pre + rank;
}

// inverse of dof_of_lattice
__host__ __device__ inline void lattice_of_dof(const GridDev& g, long dof, int* p)
{
    long tail[MAXD];
    long t = 1;
    for (int a = g.d - 1; a >= 0; a--) { tail[a] = t; t *= (long)g.k * g.n[a] + 1; }
    long rem = dof, mprod = 1;
    int c[MAXD];
    for (int a = 0; a < g.d; a++) {
        const long unit = mprod * tail[a];
        const long q = rem / unit;
        const int ca = q < g.k + 1 ? 0 : (int)((q - 1) / g.k);
        c[a] = ca;
        const long S = ca > 0 ? (long)ca * g.k + 1 : 0;
        rem -= S * unit;
        mprod *= ca > 0 ? g.k : g.k + 1;
    }
    // rem = rank inside the owner cell, axis 0 fastest
    for (int a = 0; a < g.d; a++) {
        const long m = c[a] > 0 ? g.k : g.k + 1;
        const int la = (int)(rem % m) + (c[a] > 0 ? 1 : 0);
        rem /= m;
        p[a] = c[a] * g.k + la;
    }
}

inline GridDev to_dev(const Grid& g)
{
    GridDev r;
    r.d = g.d; r.k = g.k;
    for (int a = 0; a < MAXD; a++) r.n[a] = a < g.d ? g.n[a] : 1;
    return r;
}

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
}  // namespace jfem

struct jfem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // reduction scratch: partial sums and a small result area (device)
    double* red_partials = nullptr;   // RED_BLOCKS * RED_SLOTS doubles
    double* red_result = nullptr;     // RED_SLOTS doubles (device)
    double* red_host = nullptr;       // pinned mirror of red_result
    double* scal = nullptr;           // device scalars used by fused solver kernels
    // work vectors of the CG / FCG / BiCGStab solvers: grown on demand and kept until the
    // context is destroyed, so repeated solves allocate nothing
    double* work = nullptr;
    size_t work_n = 0;                // doubles held by work
    double* noalloc = nullptr;        // this GPU's copy of the no-allocation scratch array
    int cu_count = 256;
    // set-up scratch: device temporaries of assembly and matrix set-up in a few independent
    // slots, grown on demand and kept (hipFree of a temporary costs a device-wide wait), and a
    // pinned host area for small read-backs
    unsigned char* scratch[3] = {nullptr, nullptr, nullptr};
    size_t scratch_n[3] = {0, 0, 0};
    long* pin = nullptr;              // PIN_LONGS values (pinned host memory)
    // a second stream (default priority) for set-up work that is independent of what the
    // context's stream is doing, and the event by which the context's stream waits for it
    hipStream_t aux = nullptr;
    hipEvent_t aux_done = nullptr;
};

struct jfem_space {
    jfem_ctx* ctx = nullptr;
    jfem::Grid grid;
};

struct jfem_vec {
    jfem_ctx* ctx = nullptr;
    long n = 0;
    double* d = nullptr;     // device storage (float64)
};

// CSR-stream row block: rows [r0, r1), entries [p0, p1)
struct __attribute__((aligned(16))) StBlk {
    int r0, r1, p0, p1;
};

// entries allocated past nnz in colidx / vals / vals32, so that 16-byte loads
// that start at an entry below nnz stay inside the allocation
constexpr long NNZ_PAD = 4;

struct jfem_mat {
    jfem_ctx* ctx = nullptr;
    long rows = 0, cols = 0, nnz = 0;
    int* rowptr = nullptr;   // device, rows+1 (int32: nnz < 2^31 enforced)
    int* colidx = nullptr;   // device, nnz
    double* vals = nullptr;  // device, nnz
    float* vals32 = nullptr; // optional FP32 copy of the values
    double* diag = nullptr;  // optional: diagonal (device) for Jacobi-type smoothers
    int lanes = 8;           // lanes per row for the SpMV, from the mean row length
    StBlk* blk = nullptr;    // device: CSR-stream block descriptors for FP64 values
    long nrblk = 0;
    StBlk* blk32 = nullptr;  // the same for FP32 values (twice the entries per block)
    long nrblk32 = 0;
    long blk_cap = 0, blk32_cap = 0;   // entries allocated for blk / blk32 (and part)
    double* part = nullptr;  // device: one partial sum per stream block
    jfem_mat* T = nullptr;   // cached transpose for tvMul (dropped when values change)
    // the hierarchy whose finest level operator is this matrix (set when the hierarchy is built
    // with a matrix-free finest level, cleared when the values change): the Krylov solvers
    // apply the matrix through that operator
    jfem_mg* op = nullptr;
};

namespace jfem {

constexpr int RED_BLOCKS = 2048;  // fixed reduction grid: deterministic partial-sum order
constexpr int RED_SLOTS = 8;      // simultaneous reductions per launch

inline int grid_for(long n, int block, int max_blocks = 1 << 20)
{
    long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_blocks) g = max_blocks;
    return (int)g;
}

// ----- device helpers shared by several translation units -----
void vec_fill(jfem_ctx* c, double* x, double v, long n);
void vec_copy(jfem_ctx* c, double* dst, const double* src, long n);
// y = a*x + b*y
void vec_axpby(jfem_ctx* c, long n, double a, const double* x, double b, double* y);
// dots: computes up to RED_SLOTS dot products (xs[i], ys[i]) in one pass; result on host
void vec_dots(jfem_ctx* c, long n, int m, const double* const* xs, const double* const* ys, double* out_host);
// device allocation freed when the owner leaves scope (also on a thrown Error)
template <typename T>
struct DevBuf {
    T* p = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    void alloc(size_t n)
    {
        if (p) { (void)hipFree(p); p = nullptr; }
        if (hipMalloc(&p, sizeof(T) * (n > 0 ? n : 1)) != hipSuccess) {
            p = nullptr;
            throw Error{JFEM_ERR_NOMEM, "device allocation failed"};
        }
    }
    ~DevBuf() { (void)hipFree(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    operator T*() const { return p; }
};

// a non-owning view of device memory with DevBuf's interface (.p, conversion to T*)
template <typename T>
struct DevPtr {
    T* p;
    operator T*() const { return p; }
};
// the context's solver workspace with room for n doubles (reallocated only to grow)
double* ctx_work(jfem_ctx* c, size_t n);
// set-up scratch slot `slot` (0: assembly, 1: assembly's cell tables, 2: CSR-stream block
// split) with room for `bytes`; growing it waits for the context's stream first
constexpr int SCRATCH_ASM = 0, SCRATCH_CELLS = 1, SCRATCH_SPLIT = 2;
void* ctx_scratch(jfem_ctx* c, int slot, size_t bytes);
constexpr size_t PIN_LONGS = 65536;
// 256-byte aligned offsets of consecutive arrays in one scratch slot
struct ScratchPlan {
    size_t total = 0;
    size_t add(size_t bytes)
    {
        const size_t o = total;
        total += (bytes + 255) / 256 * 256;
        return o;
    }
};
// device address of the static scratch array of the no-allocation entry points on the context's
// GPU (4 * JFEM_NOALLOC_MAX_ROWS doubles), resolved once per device by jfem_ctx_create
double* noalloc_scratch(const jfem_ctx* c);

double vec_dot(jfem_ctx* c, long n, const double* x, const double* y);
// deterministic two-stage reduction of partials already written into ctx->red_partials
void reduce_partials(jfem_ctx* c, int slots, int blocks, bool to_host, double* out_host);

// SpMV y = A x (FP64)
void spmv(jfem_mat* A, const double* x, double* y);
// SpMV y = A x with FP32 values and FP32 vectors
void spmv32(jfem_mat* A, const float* x, float* y);
// SpMV y = A x and partial dot(x, y) into reduction slot 0 (not finalised)
int spmv_dot(jfem_mat* A, const double* x, double* y);   // returns the number of partials
// y = A x with dot(x, y) reduced into the device scalar dst
void spmv_dot_dev(jfem_mat* A, const double* x, double* y, double* dst);
int choose_lanes(double mean_row);
// rowptr[0..n] (device) from per-row counts (device); returns the total
long scan_counts(jfem_ctx* c, const int* counts, long n, int* rowptr);
// the same with a caller's workspace of scan_work_bytes(n) bytes; with known_total >= 0 (the
// caller knows the sum) nothing is read back and nothing waits
size_t scan_work_bytes(long n);
long scan_counts(jfem_ctx* c, const int* counts, long n, int* rowptr, long* work, long known_total,
                 hipStream_t st = nullptr);
// lexicographic node rank (last axis fastest) of every DOF of a scalar space, into a device array
void space_lex_rank(jfem_space* s, int* out_dev);
// finest grid of a multigrid hierarchy
const Grid& mg_finest_grid(const jfem_mg* mg);
// boundary mask of the finest level (device, DOF order)
const unsigned char* mg_finest_bnd(const jfem_mg* mg);
// the stream a hierarchy's cycles run on (its context's)
hipStream_t mg_stream(const jfem_mg* mg);
// reduce partials into a device scalar (no host sync)
void reduce_partials_dev(jfem_ctx* c, int slots, int blocks, double* dst_dev);
void mat_make_f32(jfem_mat* A);
void mat_build_blocks(jfem_mat* A);
// the same without waiting: mat_blocks_queue queues the whole split (block arrays sized by a
// bound, the totals read back into ctx->pin[PIN_LONGS - 4, PIN_LONGS - 2)); mat_blocks_done
// takes the totals after the caller's next wait on the context's stream
void mat_blocks_queue(jfem_mat* A, hipStream_t st = nullptr);
void mat_blocks_done(jfem_mat* A);
// sum of n partials (device) into dst (device), deterministic
void reduce_array_dev(jfem_ctx* c, const double* partials, long n, double* dst);
void mat_diag(jfem_mat* A, double* out);
void mat_free(jfem_mat* A);
// the transpose of A (built on first use, cached until A's values change)
jfem_mat* mat_transpose(jfem_mat* A);
void mat_drop_transpose(jfem_mat* A);
void apply_dirichlet_mask(jfem_mat* A, double* rhs, const unsigned char* fixed_dev, const double* g_dev);
void space_boundary_mask(jfem_space* s, unsigned char* mask_dev);
jfem_mat* assemble_matrix(jfem_space* s, int nint, const int* kinds, const double* weights, const double* coeff_host,
                          int rule);
// 1D stiffness and mass matrices ((k+1) x (k+1), row-major) of the first cell of axis a
void element_1d(const Grid& g, int a, int rule, double* K1, double* M1);
// multigrid preconditioner application z = M r (FP64 vectors on the context stream)
void mg_apply(jfem_mg* mg, const double* r, double* z);
// y = A x through the matrix-free operator of the hierarchy A->op names (the same operator as A
// to rounding); false when A has none
bool mg_op_apply(const jfem_mat* A, const double* x, double* y);
long mg_size(const jfem_mg* mg);
// learned-smoother fallback rule: called with the residual reduction of a cycle/iteration
void mg_observe_reduction(jfem_mg* mg, double ratio);

}  // namespace jfem
