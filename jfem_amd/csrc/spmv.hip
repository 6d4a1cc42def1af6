// CSR sparse matrix-vector product and matrix handles (SparseMatrix.mvMul,
// M/linalg/SparseMatrix.java:479-489; SparseMvMul.mvMul, M/linalg/SparseMvMul.java:75-94),
// plus Dirichlet elimination (AcceptsMatrixBoundaryValues.overWriteValue,
// M/basic/AcceptsMatrixBoundaryValues.java:18-44).
//
// SpMV layout: L lanes per row (L from the mean row length), each lane
// strides the row's entries with stride L so a group reads a contiguous
// segment of values/columns; group partial sums are combined with wavefront
// shuffles.  The kernel is HBM-bound: 12 bytes per entry (8-byte value,
// 4-byte column) plus the row pointer, x (once, through L2/MALL) and y.
#include "common.hpp"
#include "stream.hpp"

#include <hipcub/hipcub.hpp>

#include <cstring>

namespace jfem {

constexpr int SPMV_BS = 256;

template <int L, typename VT, typename XT>
__global__ __launch_bounds__(SPMV_BS) void k_spmv(long nrows, const int* __restrict__ rowptr,
                                                  const int* __restrict__ cols, const VT* __restrict__ vals,
                                                  const XT* __restrict__ x, XT* __restrict__ y)
{
    const int lane = threadIdx.x % L;
    const long groups_total = (long)gridDim.x * (SPMV_BS / L);
    for (long row = (blockIdx.x * (long)SPMV_BS + threadIdx.x) / L; row < nrows; row += groups_total) {
        const int s = rowptr[row], e = rowptr[row + 1];
        XT acc = 0;
        for (int p = s + lane; p < e; p += L) acc += (XT)vals[p] * x[cols[p]];
        for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, L);
        if (lane == 0) y[row] = acc;
    }
}

// y = A x and per-block partial of dot(x, y) (reduction slot 0)
template <int L>
__global__ __launch_bounds__(SPMV_BS) void k_spmv_dot(long nrows, const int* __restrict__ rowptr,
                                                      const int* __restrict__ cols, const double* __restrict__ vals,
                                                      const double* __restrict__ x, double* __restrict__ y,
                                                      double* __restrict__ partials)
{
    __shared__ double sh[SPMV_BS / 64];
    const int lane = threadIdx.x % L;
    const long groups_total = (long)gridDim.x * (SPMV_BS / L);
    double dacc = 0;
    for (long row = (blockIdx.x * (long)SPMV_BS + threadIdx.x) / L; row < nrows; row += groups_total) {
        const int s = rowptr[row], e = rowptr[row + 1];
        double acc = 0;
        for (int p = s + lane; p < e; p += L) acc += vals[p] * x[cols[p]];
        for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, L);
        if (lane == 0) { y[row] = acc; dacc += x[row] * acc; }
    }
    for (int o = 32; o > 0; o >>= 1) dacc += __shfl_xor(dacc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = dacc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int i = 0; i < SPMV_BS / 64; i++) r += sh[i];
        partials[blockIdx.x] = r;
    }
}

static int spmv_grid(const jfem_mat* A, int L)
{
    // enough groups for one pass over the rows, capped at 16 waves per SIMD
    const long need = (A->rows * L + SPMV_BS - 1) / SPMV_BS;
    const long cap = (long)A->ctx->cu_count * 32;
    return (int)(need < 1 ? 1 : (need > cap ? cap : need));
}

template <typename VT, typename XT>
static void launch_spmv(jfem_mat* A, const VT* vals, const XT* x, XT* y)
{
    const int L = A->lanes;
    const int grid = spmv_grid(A, L);
    hipStream_t st = A->ctx->stream;
    switch (L) {
    case 4: k_spmv<4, VT, XT><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, vals, x, y); break;
    case 8: k_spmv<8, VT, XT><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, vals, x, y); break;
    case 16: k_spmv<16, VT, XT><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, vals, x, y); break;
    case 32: k_spmv<32, VT, XT><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, vals, x, y); break;
    default: k_spmv<64, VT, XT><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, vals, x, y); break;
    }
    JFEM_HIP(hipGetLastError());
}



void spmv(jfem_mat* A, const double* x, double* y)
{
    if (A->blk) launch_stream<double, double>(A, A->vals, x, EpiStore<double>{y}, nullptr);
    else launch_spmv<double, double>(A, A->vals, x, y);
}

void spmv32(jfem_mat* A, const float* x, float* y)
{
    JFEM_CHECK(A->vals32, JFEM_ERR_STATE, "matrix has no FP32 values");
    if (A->blk) launch_stream<float, float>(A, A->vals32, x, EpiStore<float>{y}, nullptr);
    else launch_spmv<float, float>(A, A->vals32, x, y);
}

// y = A x; dot(x, y) reduced into dst (device scalar)
void spmv_dot_dev(jfem_mat* A, const double* x, double* y, double* dst)
{
    JFEM_CHECK(A->blk, JFEM_ERR_STATE, "matrix has no stream blocks");
    launch_stream<double, double>(A, A->vals, x, EpiStoreDot{y, x}, A->part);
    reduce_array_dev(A->ctx, A->part, A->nrblk, dst);
}

int spmv_dot(jfem_mat* A, const double* x, double* y)
{
    const int L = A->lanes;
    int grid = spmv_grid(A, L);
    if (grid > RED_BLOCKS) grid = RED_BLOCKS;
    hipStream_t st = A->ctx->stream;
    double* part = A->ctx->red_partials;
    switch (L) {
    case 4: k_spmv_dot<4><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, A->vals, x, y, part); break;
    case 8: k_spmv_dot<8><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, A->vals, x, y, part); break;
    case 16: k_spmv_dot<16><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, A->vals, x, y, part); break;
    case 32: k_spmv_dot<32><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, A->vals, x, y, part); break;
    default: k_spmv_dot<64><<<grid, SPMV_BS, 0, st>>>(A->rows, A->rowptr, A->colidx, A->vals, x, y, part); break;
    }
    JFEM_HIP(hipGetLastError());
    return grid;
}

// Split of the rows into CSR-stream blocks: at most ST_BS rows and at most
// `cap` entries counted from the entry index rounded down to a multiple of
// `vec` (a row that alone exceeds that gets a block of its own).  The rows are
// cut into segments of ST_SEG_ROWS rows and one thread splits each segment
// greedily; pass 1 counts the blocks of each segment, pass 2 writes them at
// the segment's offset.  Blocks never cross a segment boundary, so the split
// depends only on the row pointer.
constexpr long ST_SEG_ROWS = 4096;

__global__ __launch_bounds__(64) void k_split_blocks(const int* __restrict__ rowptr, long n, long cap, int vec, long nseg,
                                                    int* __restrict__ seg_count, const long* __restrict__ seg_off,
                                                    StBlk* __restrict__ out, long maxblk)
{
    // one wavefront per segment: its row pointers are staged in LDS, then lane 0
    // walks the segment; the end of each block is found by binary search (the
    // row pointer is nondecreasing), which gives the same split as extending the
    // block one row at a time while the next row still fits
    __shared__ int rp[ST_SEG_ROWS + 1];
    const long sg = blockIdx.x;
    if (sg >= nseg) return;
    const long r0 = sg * ST_SEG_ROWS;
    const long end = r0 + ST_SEG_ROWS < n ? r0 + ST_SEG_ROWS : n;
    const int len = (int)(end - r0);
    for (int i = threadIdx.x; i <= len; i += blockDim.x) rp[i] = rowptr[r0 + i];
    __syncthreads();
    if (threadIdx.x != 0) return;
    long pos = out ? seg_off[sg] : 0;
    int cnt = 0;
    int r = 0;
    while (r < len) {
        const long a0 = (long)rp[r] & ~(long)(vec - 1);
        // smallest e in [r + 1, hi) whose next row does not fit, else hi
        const int hi = r + ST_BS < len ? r + ST_BS : len;
        int lo_e = r + 1, hi_e = hi;
        while (lo_e < hi_e) {
            const int mid = (lo_e + hi_e) >> 1;
            if ((long)rp[mid + 1] - a0 <= cap) lo_e = mid + 1;
            else hi_e = mid;
        }
        const int e = lo_e;
        if (out && pos < maxblk) out[pos] = StBlk{(int)(r0 + r), (int)(r0 + e), rp[r], rp[e]};
        pos++;
        cnt++;
        r = e;
    }
    if (!out) seg_count[sg] = cnt;
}

// exclusive prefix sums of the per-segment block counts of both splits (cnt[v nseg + i]) into
// off, and the two totals into tot; one workgroup
__global__ __launch_bounds__(256) void k_split_offsets(const int* __restrict__ cnt, long nseg, long* __restrict__ off,
                                                       long* __restrict__ tot)
{
    __shared__ long part[256];
    for (int v = 0; v < 2; v++) {
        const long per = (nseg + 255) / 256, b = threadIdx.x * per, e = b + per < nseg ? b + per : nseg;
        long s = 0;
        for (long i = b; i < e; i++) s += cnt[v * nseg + i];
        part[threadIdx.x] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            long run = 0;
            for (int t = 0; t < 256; t++) {
                const long x = part[t];
                part[t] = run;
                run += x;
            }
            tot[v] = run;
        }
        __syncthreads();
        long run = part[threadIdx.x];
        for (long i = b; i < e; i++) {
            off[v * nseg + i] = run;
            run += cnt[v * nseg + i];
        }
        __syncthreads();
    }
}

// An upper bound on the blocks of a split: a block that ends because the next row does not
// fit holds, with the first row of the next block, more than cap - vec + 1 entries, so such
// blocks number fewer than 2 nnz / (cap - vec + 1); the others end at ST_BS rows or at a
// segment's end.
static long blocks_bound(long n, long nnz, long cap, int vec, long nseg)
{
    return 2 * nnz / (cap - vec + 1) + n / ST_BS + 2 * nseg + 2;
}

// The CSR-stream row blocks of A for FP64 values (cap ST_CAP) and FP32 values (ST_CAP32),
// queued without waiting: count passes, offsets on the device, write passes into block arrays
// of the bound's size; mat_blocks_done takes the totals after the caller's next wait on the
// stream.  Counts and offsets live in the context's set-up scratch, the totals come back into
// the pinned area's tail.
void mat_blocks_queue(jfem_mat* A, hipStream_t st)
{
    jfem_ctx* c = A->ctx;
    if (!st) st = c->stream;
    (void)hipFree(A->blk);
    (void)hipFree(A->blk32);
    (void)hipFree(A->part);
    A->blk = A->blk32 = nullptr;
    A->part = nullptr;
    A->nrblk = A->nrblk32 = 0;
    const long n = A->rows;
    const long nseg = (n + ST_SEG_ROWS - 1) / ST_SEG_ROWS;
    const long caps[2] = {ST_CAP, ST_CAP32};
    const int vecs[2] = {st_vec<double>(), st_vec<float>()};
    long bound[2];
    for (int v = 0; v < 2; v++) bound[v] = blocks_bound(n, A->nnz, caps[v], vecs[v], nseg);
    A->blk_cap = bound[0];
    A->blk32_cap = bound[1];
    JFEM_HIP(hipMalloc(&A->blk, sizeof(StBlk) * bound[0]));
    JFEM_HIP(hipMalloc(&A->blk32, sizeof(StBlk) * bound[1]));
    JFEM_HIP(hipMalloc(&A->part, sizeof(double) * bound[0]));
    long* tot = c->pin + PIN_LONGS - 4;   // the pinned tail's block totals
    if (nseg == 0) {
        tot[0] = tot[1] = 0;
        return;
    }
    ScratchPlan plan;
    const size_t o_cnt = plan.add(sizeof(int) * 2 * nseg), o_off = plan.add(sizeof(long) * 2 * nseg);
    const size_t o_tot = plan.add(sizeof(long) * 2);
    unsigned char* ws = static_cast<unsigned char*>(ctx_scratch(c, SCRATCH_SPLIT, plan.total));
    int* cnt_d = reinterpret_cast<int*>(ws + o_cnt);
    long* off_d = reinterpret_cast<long*>(ws + o_off);
    long* tot_d = reinterpret_cast<long*>(ws + o_tot);
    for (int v = 0; v < 2; v++)
        k_split_blocks<<<(int)nseg, 64, 0, st>>>(A->rowptr, n, caps[v], vecs[v], nseg, cnt_d + v * nseg, nullptr,
                                                        nullptr, 0);
    k_split_offsets<<<1, 256, 0, st>>>(cnt_d, nseg, off_d, tot_d);
    k_split_blocks<<<(int)nseg, 64, 0, st>>>(A->rowptr, n, caps[0], vecs[0], nseg, nullptr, off_d, A->blk,
                                                    bound[0]);
    k_split_blocks<<<(int)nseg, 64, 0, st>>>(A->rowptr, n, caps[1], vecs[1], nseg, nullptr, off_d + nseg,
                                                    A->blk32, bound[1]);
    JFEM_HIP(hipGetLastError());
    JFEM_HIP(hipMemcpyAsync(tot, tot_d, sizeof(long) * 2, hipMemcpyDeviceToHost, st));
}

void mat_blocks_done(jfem_mat* A)
{
    const long* tot = A->ctx->pin + PIN_LONGS - 4;
    JFEM_CHECK(tot[0] <= A->blk_cap && tot[1] <= A->blk32_cap, JFEM_ERR_STATE, "CSR-stream blocks exceed their bound");
    A->nrblk = tot[0];
    A->nrblk32 = tot[1];
}

void mat_build_blocks(jfem_mat* A)
{
    mat_blocks_queue(A);
    JFEM_HIP(hipStreamSynchronize(A->ctx->stream));
    mat_blocks_done(A);
}

// ---- Dirichlet elimination, one thread group per row ----
// fixed rows: zeroed, diagonal set to 0 + (1 - 0) = 1, rhs = g.
// free rows: rhs_j += -(A_ji) * g_i for fixed columns i in ascending order,
// then those entries are zeroed (entries stay in the pattern).
// Dirichlet elimination, 16 lanes per row (coalesced over the row's entries).  A fixed row
// becomes the identity row and its right-hand side the boundary value; in a free row every
// column of a fixed node is zeroed.  With boundary values g, a free row that has such columns
// moves -A_rc g_c to its right-hand side in ascending column order, on one lane, as the
// reference's loop does (AcceptsMatrixBoundaryValues.java:18-44); rows without fixed columns
// are left alone
__global__ __launch_bounds__(256) void k_dirichlet(long n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                                   double* __restrict__ vals, const unsigned char* __restrict__ fixed,
                                                   const double* __restrict__ g, double* __restrict__ rhs,
                                                   int* __restrict__ missing_diag)
{
#pragma clang fp contract(off)
    const int sub = threadIdx.x & 15;
    const bool move = rhs && g;
    for (long r = (blockIdx.x * 256L + threadIdx.x) >> 4; r < n; r += (gridDim.x * 256L) >> 4) {
        const int s = rowptr[r], e = rowptr[r + 1];
        int found = 0;
        if (fixed[r]) {
            for (int p = s + sub; p < e; p += 16) {
                const bool dg = cols[p] == r;
                vals[p] = dg ? 1.0 : 0.0;
                found |= dg;
            }
            for (int o = 8; o > 0; o >>= 1) found |= __shfl_xor(found, o, 16);
            if (sub == 0) {
                if (!found) *missing_diag = 1;
                if (rhs) rhs[r] = g ? g[r] : 0.0;
            }
        } else {
            for (int p = s + sub; p < e; p += 16) {
                if (fixed[cols[p]]) {
                    found = 1;
                    if (!move) vals[p] = 0;
                }
            }
            if (move) {
                for (int o = 8; o > 0; o >>= 1) found |= __shfl_xor(found, o, 16);
                if (found && sub == 0) {
                    double acc = rhs[r];
                    for (int p = s; p < e; p++) {
                        const int c = cols[p];
                        if (fixed[c]) {
                            acc += -vals[p] * g[c];
                            vals[p] = 0;
                        }
                    }
                    rhs[r] = acc;
                }
            }
        }
    }
}

__global__ void k_to_f32(long n, const double* __restrict__ a, float* __restrict__ b)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (float)a[i];
}

// diagonal of a canonical CSR matrix (at most one entry per row on the diagonal), 16 lanes per row
__global__ __launch_bounds__(256) void k_diag(long n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                              const double* __restrict__ vals, double* __restrict__ diag)
{
    const int sub = threadIdx.x & 15;
    for (long r = (blockIdx.x * 256L + threadIdx.x) >> 4; r < n; r += (gridDim.x * 256L) >> 4) {
        double d = 0;
        for (int p = rowptr[r] + sub; p < rowptr[r + 1]; p += 16)
            if (cols[p] == r) d = vals[p];
        for (int o = 8; o > 0; o >>= 1) d += __shfl_xor(d, o, 16);
        if (sub == 0) diag[r] = d;
    }
}

void mat_make_f32(jfem_mat* A)
{
    if (A->vals32) return;
    JFEM_HIP(hipMalloc(&A->vals32, sizeof(float) * (A->nnz + NNZ_PAD)));
    if (A->nnz) k_to_f32<<<grid_for(A->nnz, 256, 16384), 256, 0, A->ctx->stream>>>(A->nnz, A->vals, A->vals32);
    JFEM_HIP(hipGetLastError());
}

void mat_diag(jfem_mat* A, double* out)
{
    k_diag<<<grid_for(A->rows * 16, 256, 16384), 256, 0, A->ctx->stream>>>(A->rows, A->rowptr, A->colidx, A->vals, out);
    JFEM_HIP(hipGetLastError());
}

// row length plus one for a fixed row without a stored diagonal
__global__ void k_count_with_diag(long n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                                  const unsigned char* __restrict__ fixed, int* __restrict__ counts,
                                  int* __restrict__ nmissing)
{
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
        const int s = rowptr[r], e = rowptr[r + 1];
        int add = 0;
        if (fixed[r]) {
            add = 1;
            for (int p = s; p < e; p++) if (cols[p] == r) { add = 0; break; }
        }
        counts[r] = e - s + add;
        if (add) atomicAdd(nmissing, 1);
    }
}

// copy of the rows with an explicit 0.0 diagonal entry inserted where the
// new row pointer reserves one (columns stay ascending)
__global__ void k_insert_diag(long n, const int* __restrict__ rowptr, const int* __restrict__ cols,
                              const double* __restrict__ vals, const int* __restrict__ nrowptr, int* __restrict__ ncols,
                              double* __restrict__ nvals)
{
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
        const int s = rowptr[r], e = rowptr[r + 1];
        int o = nrowptr[r];
        const bool ins = nrowptr[r + 1] - o != e - s;
        bool done = !ins;
        for (int p = s; p < e; p++) {
            if (!done && cols[p] > r) { ncols[o] = (int)r; nvals[o] = 0.0; o++; done = true; }
            ncols[o] = cols[p];
            nvals[o] = vals[p];
            o++;
        }
        if (!done) { ncols[o] = (int)r; nvals[o] = 0.0; }
    }
}

// SparseMatrix.set(1, i, i) of overWriteValue (M/basic/AcceptsMatrixBoundaryValues.java:25-43)
// creates the diagonal entry of a fixed row that has none: the rows are
// rebuilt with those entries inserted before the elimination
static void insert_fixed_diagonals(jfem_mat* A, const unsigned char* fixed_dev)
{
    jfem_ctx* c = A->ctx;
    const long n = A->rows;
    DevBuf<int> counts(n), nmiss(1);
    JFEM_HIP(hipMemsetAsync(nmiss, 0, sizeof(int), c->stream));
    k_count_with_diag<<<grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, A->rowptr, A->colidx, fixed_dev, counts,
                                                                       nmiss);
    JFEM_HIP(hipGetLastError());
    int h = 0;
    JFEM_HIP(hipMemcpyAsync(&h, nmiss, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    JFEM_HIP(hipStreamSynchronize(c->stream));
    if (h == 0) return;
    int* nrowptr = nullptr;
    int* ncols = nullptr;
    double* nvals = nullptr;
    try {
        JFEM_HIP(hipMalloc(&nrowptr, sizeof(int) * (n + 1)));
        const long nnz = scan_counts(c, counts, n, nrowptr);
        JFEM_HIP(hipMalloc(&ncols, sizeof(int) * (nnz + NNZ_PAD)));
        JFEM_HIP(hipMalloc(&nvals, sizeof(double) * (nnz + NNZ_PAD)));
        JFEM_HIP(hipMemsetAsync(ncols, 0, sizeof(int) * (nnz + NNZ_PAD), c->stream));
        JFEM_HIP(hipMemsetAsync(nvals, 0, sizeof(double) * (nnz + NNZ_PAD), c->stream));
        k_insert_diag<<<grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, A->rowptr, A->colidx, A->vals, nrowptr, ncols,
                                                                       nvals);
        JFEM_HIP(hipGetLastError());
        JFEM_HIP(hipStreamSynchronize(c->stream));
        (void)hipFree(A->rowptr);
        (void)hipFree(A->colidx);
        (void)hipFree(A->vals);
        A->rowptr = nrowptr;
        A->colidx = ncols;
        A->vals = nvals;
        nrowptr = nullptr;
        ncols = nullptr;
        nvals = nullptr;
        A->nnz = nnz;
    } catch (...) {
        (void)hipFree(nrowptr);
        (void)hipFree(ncols);
        (void)hipFree(nvals);
        throw;
    }
    (void)hipFree(A->diag);
    A->diag = nullptr;
    A->lanes = choose_lanes(n > 0 ? (double)A->nnz / n : 1.0);
    mat_build_blocks(A);
    if (A->vals32) {
        (void)hipFree(A->vals32);
        A->vals32 = nullptr;
        mat_make_f32(A);
    }
}

void apply_dirichlet_mask(jfem_mat* A, double* rhs, const unsigned char* fixed_dev, const double* g_dev)
{
    jfem_ctx* c = A->ctx;
    mat_drop_transpose(A);   // the values change
    A->op = nullptr;
    insert_fixed_diagonals(A, fixed_dev);
    int* flag = nullptr;
    JFEM_HIP(hipMalloc(&flag, sizeof(int)));
    JFEM_HIP(hipMemsetAsync(flag, 0, sizeof(int), c->stream));
    k_dirichlet<<<grid_for(A->rows * 16, 256, 16384), 256, 0, c->stream>>>(A->rows, A->rowptr, A->colidx, A->vals,
                                                                        fixed_dev, g_dev, rhs, flag);
    int h = 0;
    hipError_t e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    hipError_t e2 = hipStreamSynchronize(c->stream);
    (void)hipFree(flag);
    JFEM_HIP(e);
    JFEM_HIP(e2);
    JFEM_CHECK(!h, JFEM_ERR_STATE, "Dirichlet node without a diagonal entry in the pattern");
    if (A->vals32) {
        (void)hipFree(A->vals32);
        A->vals32 = nullptr;
        mat_make_f32(A);
    }
}

// ---- transpose (SparseMatrix.tvMul, M/linalg/SparseMatrix.java:491-501) ----
__global__ void k_entry_rows(long n, const int* __restrict__ rowptr, int* __restrict__ rows)
{
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
        for (int p = rowptr[r]; p < rowptr[r + 1]; p++) rows[p] = (int)r;
}

__global__ void k_iota(long m, int* __restrict__ v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m; i += (long)gridDim.x * blockDim.x) v[i] = (int)i;
}

__global__ void k_col_counts(long m, const int* __restrict__ cols, int* __restrict__ counts)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m; i += (long)gridDim.x * blockDim.x)
        atomicAdd(&counts[cols[i]], 1);
}

__global__ void k_gather_transpose(long m, const int* __restrict__ perm, const int* __restrict__ rows,
                                   const double* __restrict__ vals, int* __restrict__ tcols, double* __restrict__ tvals)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m; i += (long)gridDim.x * blockDim.x) {
        const int e = perm[i];
        tcols[i] = rows[e];
        tvals[i] = vals[e];
    }
}

void mat_drop_transpose(jfem_mat* A)
{
    if (A && A->T) {
        mat_free(A->T);
        A->T = nullptr;
    }
}

// entries sorted by column with a stable radix sort (equal columns keep their row order),
// so every row of the transpose lists its columns ascending and the result is deterministic
jfem_mat* mat_transpose(jfem_mat* A)
{
    if (A->T) return A->T;
    jfem_ctx* c = A->ctx;
    hipStream_t st = c->stream;
    const long m = A->nnz;
    auto* T = new jfem_mat();
    T->ctx = c;
    T->rows = A->cols;
    T->cols = A->rows;
    T->nnz = m;
    try {
        DevBuf<int> rows(m), keys_out(m), idx(m), perm(m), counts(T->rows);
        JFEM_HIP(hipMalloc(&T->rowptr, sizeof(int) * (T->rows + 1)));
        JFEM_HIP(hipMalloc(&T->colidx, sizeof(int) * (m + NNZ_PAD)));
        JFEM_HIP(hipMalloc(&T->vals, sizeof(double) * (m + NNZ_PAD)));
        JFEM_HIP(hipMemsetAsync(T->colidx, 0, sizeof(int) * (m + NNZ_PAD), st));
        JFEM_HIP(hipMemsetAsync(T->vals, 0, sizeof(double) * (m + NNZ_PAD), st));
        JFEM_HIP(hipMemsetAsync(counts, 0, sizeof(int) * (T->rows > 0 ? T->rows : 1), st));
        if (A->rows) k_entry_rows<<<grid_for(A->rows, 256, 16384), 256, 0, st>>>(A->rows, A->rowptr, rows);
        if (m) {
            k_iota<<<grid_for(m, 256, 16384), 256, 0, st>>>(m, idx);
            k_col_counts<<<grid_for(m, 256, 16384), 256, 0, st>>>(m, A->colidx, counts);
            int bits = 1;
            while (bits < 31 && (1l << bits) < T->rows) bits++;
            size_t tmp_bytes = 0;
            JFEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, A->colidx, keys_out.p, idx.p, perm.p,
                                                        (int)m, 0, bits, st));
            DevBuf<char> tmp(tmp_bytes);
            JFEM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, A->colidx, keys_out.p, idx.p, perm.p,
                                                        (int)m, 0, bits, st));
            k_gather_transpose<<<grid_for(m, 256, 16384), 256, 0, st>>>(m, perm, rows, A->vals, T->colidx, T->vals);
        }
        JFEM_HIP(hipGetLastError());
        const long tot = scan_counts(c, counts, T->rows, T->rowptr);
        JFEM_CHECK(tot == m, JFEM_ERR_STATE, "transpose lost entries");
        T->lanes = choose_lanes(T->rows > 0 ? (double)m / T->rows : 1.0);
        mat_build_blocks(T);
    } catch (...) {
        mat_free(T);
        throw;
    }
    A->T = T;
    return T;
}

void mat_free(jfem_mat* A)
{
    if (!A) return;
    mat_drop_transpose(A);
    (void)hipStreamSynchronize(A->ctx->stream);
    (void)hipFree(A->rowptr);
    (void)hipFree(A->colidx);
    (void)hipFree(A->vals);
    (void)hipFree(A->vals32);
    (void)hipFree(A->diag);
    (void)hipFree(A->blk);
    (void)hipFree(A->blk32);
    (void)hipFree(A->part);
    delete A;
}

}  // namespace jfem

using namespace jfem;

extern "C" {

int jfem_mat_from_csr(jfem_ctx* ctx, long rows, long cols, const long* rowptr, const int* colidx, const double* vals,
                      jfem_mat** out)
{
    API_BEGIN
    JFEM_CHECK(ctx && rowptr && out && rows >= 0 && cols >= 0, JFEM_ERR_ARG, "null argument");
    // a CSR row pointer starts at 0 and never decreases; then every entry belongs to a row and
    // the kernels' reads stay inside colidx / vals
    JFEM_CHECK(rowptr[0] == 0, JFEM_ERR_ARG, "row pointer does not start at 0");
    const long nnz = rowptr[rows];
    JFEM_CHECK(nnz >= 0 && nnz < (1l << 31) - 1, JFEM_ERR_UNSUPPORTED, "2^31 or more entries");
    JFEM_CHECK(nnz == 0 || (colidx && vals), JFEM_ERR_ARG, "null argument");
    std::vector<int> rp(rows + 1);
    for (long i = 0; i <= rows; i++) {
        JFEM_CHECK(i == 0 || rowptr[i] >= rowptr[i - 1], JFEM_ERR_ARG, "row pointer not monotone");
        rp[i] = (int)rowptr[i];
    }
    for (long p = 0; p < nnz; p++) JFEM_CHECK(colidx[p] >= 0 && colidx[p] < cols, JFEM_ERR_ARG, "x coordinate out of bounds");
    auto* A = new jfem_mat();
    A->ctx = ctx;
    A->rows = rows;
    A->cols = cols;
    A->nnz = nnz;
    JFEM_HIP(hipMalloc(&A->rowptr, sizeof(int) * (rows + 1)));
    JFEM_HIP(hipMalloc(&A->colidx, sizeof(int) * (nnz + NNZ_PAD)));
    JFEM_HIP(hipMalloc(&A->vals, sizeof(double) * (nnz + NNZ_PAD)));
    JFEM_HIP(hipMemcpy(A->rowptr, rp.data(), sizeof(int) * (rows + 1), hipMemcpyHostToDevice));
    if (nnz) {
        JFEM_HIP(hipMemcpy(A->colidx, colidx, sizeof(int) * nnz, hipMemcpyHostToDevice));
        JFEM_HIP(hipMemcpy(A->vals, vals, sizeof(double) * nnz, hipMemcpyHostToDevice));
    }
    A->lanes = choose_lanes(rows > 0 ? (double)nnz / rows : 1.0);
    mat_build_blocks(A);
    *out = A;
    API_END
}

int jfem_mat_destroy(jfem_mat* A)
{
    API_BEGIN
    mat_free(A);
    API_END
}

int jfem_mat_info(const jfem_mat* A, long* rows, long* cols, long* nnz)
{
    API_BEGIN
    JFEM_CHECK(A, JFEM_ERR_ARG, "null matrix");
    if (rows) *rows = A->rows;
    if (cols) *cols = A->cols;
    if (nnz) *nnz = A->nnz;
    API_END
}

int jfem_mat_export_csr(const jfem_mat* A, long* rowptr, int* colidx, double* vals)
{
    API_BEGIN
    JFEM_CHECK(A && rowptr && colidx && vals, JFEM_ERR_ARG, "null argument");
    std::vector<int> rp(A->rows + 1);
    hipStream_t st = A->ctx->stream;
    JFEM_HIP(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int) * (A->rows + 1), hipMemcpyDeviceToHost, st));
    if (A->nnz) {
        JFEM_HIP(hipMemcpyAsync(colidx, A->colidx, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, st));
        JFEM_HIP(hipMemcpyAsync(vals, A->vals, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, st));
    }
    JFEM_HIP(hipStreamSynchronize(st));
    for (long i = 0; i <= A->rows; i++) rowptr[i] = rp[i];
    API_END
}

int jfem_apply_dirichlet(jfem_mat* A, jfem_vec* rhs, long nb, const int* nodes, const double* values)
{
    API_BEGIN
    JFEM_CHECK(A && (nodes || nb == 0), JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(A->rows == A->cols, JFEM_ERR_ARG, "Dirichlet elimination needs a square matrix");
    JFEM_CHECK(!rhs || rhs->n == A->rows, JFEM_ERR_ARG, "Incompatible sizes");
    const long n = A->rows;
    std::vector<unsigned char> mask(n, 0);
    std::vector<double> g;
    if (values) g.assign(n, 0.0);
    for (long b = 0; b < nb; b++) {
        JFEM_CHECK(nodes[b] >= 0 && nodes[b] < n, JFEM_ERR_ARG, "boundary node out of range");
        mask[nodes[b]] = 1;
        if (values) g[nodes[b]] = values[b];
    }
    DevBuf<unsigned char> dm(n);
    DevBuf<double> dg;
    if (values) dg.alloc(n);
    if (n) JFEM_HIP(hipMemcpyAsync(dm, mask.data(), n, hipMemcpyHostToDevice, A->ctx->stream));
    if (values) JFEM_HIP(hipMemcpyAsync(dg, g.data(), sizeof(double) * n, hipMemcpyHostToDevice, A->ctx->stream));
    // the host arrays must outlive the copies
    JFEM_HIP(hipStreamSynchronize(A->ctx->stream));
    apply_dirichlet_mask(A, rhs ? rhs->d : nullptr, dm, values ? dg.p : nullptr);
    API_END
}

int jfem_spmv_t(jfem_mat* A, const jfem_vec* x, jfem_vec* y)
{
    API_BEGIN
    JFEM_CHECK(A && x && y, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(x->n == A->rows && y->n == A->cols, JFEM_ERR_ARG, "Incompatible sizes");
    // the product reads x from every workgroup while others write y: in place it is wrong
    JFEM_CHECK(x->d != y->d || A->rows == 0 || A->cols == 0, JFEM_ERR_ARG, "tvMul output aliases its input");
    JFEM_CHECK(A->nnz < (1l << 31) - 1, JFEM_ERR_UNSUPPORTED, "2^31 or more entries");
    spmv(mat_transpose(A), x->d, y->d);
    JFEM_HIP(hipGetLastError());
    JFEM_HIP(hipStreamSynchronize(A->ctx->stream));
    API_END
}

int jfem_spmv(jfem_mat* A, const jfem_vec* x, jfem_vec* y)
{
    API_BEGIN
    JFEM_CHECK(A && x && y, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(x->n == A->cols && y->n == A->rows, JFEM_ERR_ARG, "Incompatible sizes");
    JFEM_CHECK(x->d != y->d || A->rows == 0, JFEM_ERR_ARG, "SpMV output aliases its input");
    spmv(A, x->d, y->d);
    JFEM_HIP(hipStreamSynchronize(A->ctx->stream));
    API_END
}

int jfem_spmv_f32(jfem_mat* A, const float* x, float* y)
{
    API_BEGIN
    JFEM_CHECK(A && x && y, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK((const void*)x != (const void*)y || A->rows == 0, JFEM_ERR_ARG, "SpMV output aliases its input");
    mat_make_f32(A);
    spmv32(A, x, y);
    JFEM_HIP(hipStreamSynchronize(A->ctx->stream));
    API_END
}

int jfem_spmv_bench(jfem_mat* A, const jfem_vec* x, jfem_vec* y, int reps, int fp32, double* ms)
{
    API_BEGIN
    JFEM_CHECK(A && x && y && ms && reps > 0, JFEM_ERR_ARG, "bad argument");
    JFEM_CHECK(x->n == A->cols && y->n == A->rows, JFEM_ERR_ARG, "Incompatible sizes");
    hipStream_t st = A->ctx->stream;
    float *x32 = nullptr, *y32 = nullptr;
    if (fp32) {
        mat_make_f32(A);
        JFEM_HIP(hipMalloc(&x32, sizeof(float) * A->cols));
        JFEM_HIP(hipMalloc(&y32, sizeof(float) * A->rows));
        k_to_f32<<<grid_for(A->cols, 256, 16384), 256, 0, st>>>(A->cols, x->d, x32);
    }
    hipEvent_t e0, e1;
    JFEM_HIP(hipEventCreate(&e0));
    JFEM_HIP(hipEventCreate(&e1));
    // one warm-up launch
    if (fp32) spmv32(A, x32, y32); else spmv(A, x->d, y->d);
    JFEM_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) {
        if (fp32) spmv32(A, x32, y32); else spmv(A, x->d, y->d);
    }
    JFEM_HIP(hipEventRecord(e1, st));
    JFEM_HIP(hipEventSynchronize(e1));
    float f = 0;
    JFEM_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(x32);
    (void)hipFree(y32);
    API_END
}

}  // extern "C"
