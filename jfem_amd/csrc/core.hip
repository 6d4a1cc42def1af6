// Context, vectors and BLAS-1 kernels (DenseVector, M/linalg/DenseVector.java:103-260).
// Reductions are deterministic: a fixed grid of RED_BLOCKS blocks writes one
// partial per block and a single block sums the partials in a fixed order.
#include "common.hpp"

#include <cstring>
#include <mutex>

namespace jfem {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

constexpr int BS = 256;

__global__ void k_fill(double* __restrict__ x, double v, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}

// plain copy b = a with 16-byte accesses, four in flight per thread (HBM calibration kernel)
__global__ __launch_bounds__(256) void k_copy(long n, const double* __restrict__ a, double* __restrict__ b)
{
    const long n2 = n / 2;
    const double2* __restrict__ a2 = reinterpret_cast<const double2*>(a);
    double2* __restrict__ b2 = reinterpret_cast<double2*>(b);
    const long stride = (long)gridDim.x * blockDim.x;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n2; i += 4 * stride) {
        const double2 v0 = a2[i], v1 = a2[i + stride], v2 = a2[i + 2 * stride], v3 = a2[i + 3 * stride];
        b2[i] = v0; b2[i + stride] = v1; b2[i + 2 * stride] = v2; b2[i + 3 * stride] = v3;
    }
    for (; i < n2; i += stride) b2[i] = a2[i];
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) b[n - 1] = a[n - 1];
}

// streaming read of n doubles (calibration of the achievable read bandwidth):
// each workgroup reads one 16 KiB chunk with four 16-byte loads per thread; the
// sum is stored only if it equals a value no input here produces, so the loads
// cannot be removed
__global__ __launch_bounds__(256) void k_read(long n, const double* __restrict__ a, double* __restrict__ sink)
{
    const long n2 = n / 2;
    const double2* __restrict__ a2 = reinterpret_cast<const double2*>(a);
    const long base = (long)blockIdx.x * 1024 + threadIdx.x;
    double acc = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const long i = base + u * 256;
        if (i < n2) { const double2 v = a2[i]; acc += v.x + v.y; }
    }
    if (acc == -1.2345e300) sink[0] = acc;
}

__global__ void k_axpby(long n, double a, const double* __restrict__ x, double b, double* __restrict__ y)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}

// block-wide sum of one value per thread (wavefront shuffles then LDS)
template <int NT>
__device__ inline double block_sum(double v, double* sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NT / 64; i++) r += sh[i];
    __syncthreads();
    return r;   // valid in thread 0
}

struct DotArgs {
    const double* x[RED_SLOTS];
    const double* y[RED_SLOTS];
};

__global__ __launch_bounds__(BS) void k_dots(long n, int m, DotArgs args, double* __restrict__ partials)
{
    __shared__ double sh[BS / 64];
    for (int s = 0; s < m; s++) {
        const double* __restrict__ x = args.x[s];
        const double* __restrict__ y = args.y[s];
        double acc = 0;
        for (long i = blockIdx.x * (long)BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) acc += x[i] * y[i];
        const double r = block_sum<BS>(acc, sh);
        if (threadIdx.x == 0) partials[s * RED_BLOCKS + blockIdx.x] = r;
    }
}

// sums RED_BLOCKS partials of each slot (blocks beyond `blocks` are ignored)
__global__ __launch_bounds__(BS) void k_reduce_partials(int slots, int blocks, const double* __restrict__ partials,
                                                        double* __restrict__ out)
{
    __shared__ double sh[BS / 64];
    for (int s = 0; s < slots; s++) {
        double acc = 0;
        for (int i = threadIdx.x; i < blocks; i += BS) acc += partials[s * RED_BLOCKS + i];
        const double r = block_sum<BS>(acc, sh);
        if (threadIdx.x == 0) out[s] = r;
    }
}

void vec_fill(jfem_ctx* c, double* x, double v, long n)
{
    if (n <= 0) return;
    k_fill<<<grid_for(n, BS, 8192), BS, 0, c->stream>>>(x, v, n);
    JFEM_HIP(hipGetLastError());
}

void vec_copy(jfem_ctx* c, double* dst, const double* src, long n)
{
    if (n <= 0 || dst == src) return;
    JFEM_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
}

void vec_axpby(jfem_ctx* c, long n, double a, const double* x, double b, double* y)
{
    if (n <= 0) return;
    k_axpby<<<grid_for(n, BS, 8192), BS, 0, c->stream>>>(n, a, x, b, y);
    JFEM_HIP(hipGetLastError());
}

void reduce_partials(jfem_ctx* c, int slots, int blocks, bool to_host, double* out_host)
{
    k_reduce_partials<<<1, BS, 0, c->stream>>>(slots, blocks, c->red_partials, c->red_result);
    JFEM_HIP(hipGetLastError());
    if (to_host) {
        JFEM_HIP(hipMemcpyAsync(c->red_host, c->red_result, sizeof(double) * slots, hipMemcpyDeviceToHost, c->stream));
        JFEM_HIP(hipStreamSynchronize(c->stream));
        for (int s = 0; s < slots; s++) out_host[s] = c->red_host[s];
    }
}

// arbitrary-length partial array -> one value: RED_BLOCKS blocks write fixed-order
// partials, then a single block sums them (deterministic)
__global__ __launch_bounds__(BS) void k_reduce_array(const double* __restrict__ a, long n, double* __restrict__ partials)
{
    __shared__ double sh[BS / 64];
    double acc = 0;
    for (long i = blockIdx.x * (long)BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) acc += a[i];
    const double r = block_sum<BS>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

void reduce_array_dev(jfem_ctx* c, const double* partials, long n, double* dst)
{
    const int blocks = grid_for(n, BS * 4, 256);
    k_reduce_array<<<blocks, BS, 0, c->stream>>>(partials, n, c->red_partials);
    JFEM_HIP(hipGetLastError());
    k_reduce_partials<<<1, BS, 0, c->stream>>>(1, blocks, c->red_partials, dst);
    JFEM_HIP(hipGetLastError());
}

void reduce_partials_dev(jfem_ctx* c, int slots, int blocks, double* dst_dev)
{
    k_reduce_partials<<<1, BS, 0, c->stream>>>(slots, blocks, c->red_partials, dst_dev);
    JFEM_HIP(hipGetLastError());
}

void vec_dots(jfem_ctx* c, long n, int m, const double* const* xs, const double* const* ys, double* out_host)
{
    JFEM_CHECK(m >= 1 && m <= RED_SLOTS, JFEM_ERR_ARG, "too many simultaneous dot products");
    DotArgs a;
    for (int s = 0; s < m; s++) { a.x[s] = xs[s]; a.y[s] = ys[s]; }
    const int blocks = grid_for(n, BS, RED_BLOCKS);
    k_dots<<<blocks, BS, 0, c->stream>>>(n, m, a, c->red_partials);
    JFEM_HIP(hipGetLastError());
    reduce_partials(c, m, blocks, true, out_host);
}

double vec_dot(jfem_ctx* c, long n, const double* x, const double* y)
{
    double r = 0;
    vec_dots(c, n, 1, &x, &y, &r);
    return r;
}

}  // namespace jfem

using namespace jfem;


namespace jfem {

// Scratch of the no-allocation entry points (jfem_smooth_noalloc): a file-scope device array
// that every GPU the library runs on holds until the process ends.  In host code its name is
// a host-side stand-in, not a device address; the address of each GPU's copy comes from
// hipGetSymbolAddress with that GPU current (scratch_address) and only that address is used.
static __device__ __align__(256) double scratch[4 * (size_t)JFEM_NOALLOC_MAX_ROWS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// once per device, with `device` current
double* scratch_address(int device)
{
    JFEM_CHECK(device >= 0 && device < kMaxDevices, JFEM_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* ptr = nullptr;
        JFEM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        JFEM_CHECK(ptr != nullptr, JFEM_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[device] = static_cast<double*>(ptr);
    }
    return g_scratch_dev[device];
}
}  // namespace

double* noalloc_scratch(const jfem_ctx* c)
{
    JFEM_CHECK(c->noalloc, JFEM_ERR_STATE, "no-allocation scratch not resolved for this context");
    return c->noalloc;
}

double* ctx_work(jfem_ctx* c, size_t n)
{
    if (n > c->work_n) {
        // the old buffer may still be read by work queued on the stream
        JFEM_HIP(hipStreamSynchronize(c->stream));
        (void)hipFree(c->work);
        c->work = nullptr;
        c->work_n = 0;
        JFEM_HIP(hipMalloc(&c->work, sizeof(double) * n));
        c->work_n = n;
    }
    return c->work;
}

void* ctx_scratch(jfem_ctx* c, int slot, size_t bytes)
{
    JFEM_CHECK(slot >= 0 && slot < 3, JFEM_ERR_ARG, "scratch slot out of range");
    if (bytes > c->scratch_n[slot]) {
        // the old buffer may still be read by work queued on either stream
        JFEM_HIP(hipStreamSynchronize(c->stream));
        JFEM_HIP(hipStreamSynchronize(c->aux));
        (void)hipFree(c->scratch[slot]);
        c->scratch[slot] = nullptr;
        c->scratch_n[slot] = 0;
        JFEM_HIP(hipMalloc(&c->scratch[slot], bytes));
        c->scratch_n[slot] = bytes;
    }
    return c->scratch[slot];
}

}  // namespace jfem

extern "C" {

const char* jfem_last_error(void) { return g_last_error.c_str(); }

int jfem_ctx_create(int device, jfem_ctx** out)
{
    API_BEGIN
    JFEM_CHECK(out != nullptr, JFEM_ERR_ARG, "out is null");
    int ndev = 0;
    JFEM_HIP(hipGetDeviceCount(&ndev));
    JFEM_CHECK(device >= 0 && device < ndev, JFEM_ERR_ARG, "no such device");
    JFEM_HIP(hipSetDevice(device));
    auto* c = new jfem_ctx();
    c->device = device;
    JFEM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    JFEM_HIP(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
    JFEM_HIP(hipEventCreateWithFlags(&c->aux_done, hipEventDisableTiming));
    JFEM_HIP(hipMalloc(&c->red_partials, sizeof(double) * RED_BLOCKS * RED_SLOTS));
    JFEM_HIP(hipMalloc(&c->red_result, sizeof(double) * RED_SLOTS));
    JFEM_HIP(hipMalloc(&c->scal, sizeof(double) * 64));
    JFEM_HIP(hipHostMalloc(&c->red_host, sizeof(double) * RED_SLOTS));
    JFEM_HIP(hipHostMalloc(&c->pin, sizeof(long) * PIN_LONGS));
    hipDeviceProp_t prop;
    JFEM_HIP(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    c->noalloc = scratch_address(device);   // device is current (hipSetDevice above)
    *out = c;
    API_END
}

int jfem_ctx_destroy(jfem_ctx* c)
{
    API_BEGIN
    if (!c) return JFEM_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->aux);
    (void)hipFree(c->red_partials);
    (void)hipFree(c->red_result);
    (void)hipFree(c->scal);
    (void)hipFree(c->work);
    for (unsigned char* p : c->scratch) (void)hipFree(p);
    (void)hipHostFree(c->red_host);
    (void)hipHostFree(c->pin);
    (void)hipEventDestroy(c->aux_done);
    (void)hipStreamDestroy(c->aux);
    (void)hipStreamDestroy(c->stream);
    delete c;
    API_END
}

int jfem_ctx_sync(jfem_ctx* c)
{
    API_BEGIN
    JFEM_CHECK(c, JFEM_ERR_ARG, "null context");
    JFEM_HIP(hipStreamSynchronize(c->stream));
    API_END
}

static thread_local hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

int jfem_timer_start(jfem_ctx* c)
{
    API_BEGIN
    if (!g_t0) { JFEM_HIP(hipEventCreate(&g_t0)); JFEM_HIP(hipEventCreate(&g_t1)); }
    JFEM_HIP(hipEventRecord(g_t0, c->stream));
    API_END
}

int jfem_timer_stop(jfem_ctx* c, double* ms)
{
    API_BEGIN
    JFEM_CHECK(g_t0, JFEM_ERR_STATE, "timer not started");
    JFEM_HIP(hipEventRecord(g_t1, c->stream));
    JFEM_HIP(hipEventSynchronize(g_t1));
    float f = 0;
    JFEM_HIP(hipEventElapsedTime(&f, g_t0, g_t1));
    *ms = f;
    API_END
}

int jfem_vec_create(jfem_ctx* c, long n, jfem_vec** out)
{
    API_BEGIN
    JFEM_CHECK(c && out && n >= 0, JFEM_ERR_ARG, "bad vector arguments");
    auto* v = new jfem_vec();
    v->ctx = c;
    v->n = n;
    if (n > 0) {
        hipError_t e = hipMalloc(&v->d, sizeof(double) * n);
        if (e != hipSuccess) { delete v; throw Error{JFEM_ERR_NOMEM, "device allocation failed"}; }
        JFEM_HIP(hipMemsetAsync(v->d, 0, sizeof(double) * n, c->stream));
    }
    *out = v;
    API_END
}

int jfem_vec_destroy(jfem_vec* v)
{
    API_BEGIN
    if (!v) return JFEM_OK;
    (void)hipStreamSynchronize(v->ctx->stream);
    (void)hipFree(v->d);
    delete v;
    API_END
}

int jfem_vec_upload(jfem_vec* v, const double* host)
{
    API_BEGIN
    JFEM_CHECK(v && (host || v->n == 0), JFEM_ERR_ARG, "null argument");
    if (v->n) JFEM_HIP(hipMemcpyAsync(v->d, host, sizeof(double) * v->n, hipMemcpyHostToDevice, v->ctx->stream));
    JFEM_HIP(hipStreamSynchronize(v->ctx->stream));
    API_END
}

int jfem_vec_download(const jfem_vec* v, double* host)
{
    API_BEGIN
    JFEM_CHECK(v && (host || v->n == 0), JFEM_ERR_ARG, "null argument");
    if (v->n) JFEM_HIP(hipMemcpyAsync(host, v->d, sizeof(double) * v->n, hipMemcpyDeviceToHost, v->ctx->stream));
    JFEM_HIP(hipStreamSynchronize(v->ctx->stream));
    API_END
}

int jfem_vec_fill(jfem_vec* v, double value)
{
    API_BEGIN
    JFEM_CHECK(v, JFEM_ERR_ARG, "null vector");
    vec_fill(v->ctx, v->d, value, v->n);
    API_END
}

int jfem_vec_dot(const jfem_vec* x, const jfem_vec* y, double* out)
{
    API_BEGIN
    JFEM_CHECK(x && y && out, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(x->n == y->n, JFEM_ERR_ARG, "Vectors have different size");
    *out = vec_dot(x->ctx, x->n, x->d, y->d);
    API_END
}

int jfem_copy_bench(jfem_ctx* c, long n, int reps, double* ms)
{
    API_BEGIN
    JFEM_CHECK(c && ms && n > 0 && reps > 0, JFEM_ERR_ARG, "bad argument");
    double *a = nullptr, *b = nullptr;
    JFEM_HIP(hipMalloc(&a, sizeof(double) * n));
    JFEM_HIP(hipMalloc(&b, sizeof(double) * n));
    vec_fill(c, a, 1.0, n);
    vec_fill(c, b, 0.0, n);
    hipEvent_t e0, e1;
    JFEM_HIP(hipEventCreate(&e0));
    JFEM_HIP(hipEventCreate(&e1));
    const int g = grid_for(n / 2, BS, c->cu_count * 8);
    k_copy<<<g, BS, 0, c->stream>>>(n, a, b);
    JFEM_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) k_copy<<<g, BS, 0, c->stream>>>(n, a, b);
    JFEM_HIP(hipEventRecord(e1, c->stream));
    JFEM_HIP(hipEventSynchronize(e1));
    float f = 0;
    JFEM_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    API_END
}

int jfem_read_bench(jfem_ctx* c, long n, int reps, double* ms)
{
    API_BEGIN
    JFEM_CHECK(c && ms && n > 0 && reps > 0, JFEM_ERR_ARG, "bad argument");
    double *a = nullptr, *sink = nullptr;
    JFEM_HIP(hipMalloc(&a, sizeof(double) * n));
    JFEM_HIP(hipMalloc(&sink, sizeof(double)));
    vec_fill(c, a, 1.0, n);
    hipEvent_t e0, e1;
    JFEM_HIP(hipEventCreate(&e0));
    JFEM_HIP(hipEventCreate(&e1));
    const long g = (n / 2 + 1023) / 1024;
    JFEM_CHECK(g < (1l << 31), JFEM_ERR_ARG, "n too large");
    k_read<<<(unsigned)g, 256, 0, c->stream>>>(n, a, sink);
    JFEM_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) k_read<<<(unsigned)g, 256, 0, c->stream>>>(n, a, sink);
    JFEM_HIP(hipEventRecord(e1, c->stream));
    JFEM_HIP(hipEventSynchronize(e1));
    float f = 0;
    JFEM_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(sink);
    API_END
}

int jfem_vec_axpby(double a, const jfem_vec* x, double b, jfem_vec* y)
{
    API_BEGIN
    JFEM_CHECK(x && y, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(x->n == y->n, JFEM_ERR_ARG, "Vectors have different size");
    vec_axpby(x->ctx, x->n, a, x->d, b, y->d);
    API_END
}

int jfem_vec_copy(const jfem_vec* src, jfem_vec* dst)
{
    API_BEGIN
    JFEM_CHECK(src && dst, JFEM_ERR_ARG, "null argument");
    JFEM_CHECK(src->n == dst->n, JFEM_ERR_ARG, "Vectors have different size");
    if (src->n > 0 && src->d != dst->d)
        JFEM_HIP(hipMemcpyAsync(dst->d, src->d, sizeof(double) * src->n, hipMemcpyDeviceToDevice, src->ctx->stream));
    API_END
}

}  // extern "C"
