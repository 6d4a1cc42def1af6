// Learned smoother network on the structured node lattice: four 3^d-tap
// convolutions (cin -> 32 -> 32 -> 32 -> 1, ReLU between), trained on the GPU.
//
// Layout: pixels are lattice nodes in lattice order (axis 0 fastest); hidden
// activations are NHWC bf16 (32 channels, 64 bytes per pixel).  The two dense
// 32 -> 32 layers (forward, input gradient, weight gradient) run on MFMA
// (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32 accumulation); the thin first
// and last layers (cin and 1 channels) run as FP32 FMA kernels.  Master
// weights and Adam state are FP32; the MFMA layers read BF16 copies refreshed
// after every optimiser step.
#include "common.hpp"
#include "nn.hpp"

#include <cmath>
#include <cstring>
#include <fstream>

namespace jfem {

typedef __bf16 bfv8 __attribute__((ext_vector_type(8)));   // the same bits as bf16 for the dot-product builtins
typedef __bf16 bfv2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// lattice coordinates of pixel p (32-bit arithmetic: np < 2^31 is checked in
// make_lat); unused axes have extent 1 and coordinate 0
__device__ inline void lat_xyz(const Lat& L, long p, int* i)
{
    const unsigned u = (unsigned)p;
    const unsigned r = u / (unsigned)L.n[0];
    i[0] = (int)(u - r * (unsigned)L.n[0]);
    i[1] = (int)(r % (unsigned)L.n[1]);
    i[2] = (int)(r / (unsigned)L.n[1]);
}

// neighbour of pixel p (coordinates i) at tap t: offset ((t / 3^a) % 3) - 1 on
// axis a (axis 0 fastest in t), times sign (+1 forward conv, -1 its transpose);
// -1 if outside the lattice
__device__ inline long lat_nb(const Lat& L, long p, const int* i, int t, int sign)
{
    const int o0 = sign * (t % 3 - 1);
    const int o1 = L.d > 1 ? sign * ((t / 3) % 3 - 1) : 0;
    const int o2 = L.d > 2 ? sign * ((t / 9) % 3 - 1) : 0;
    if ((unsigned)(i[0] + o0) >= (unsigned)L.n[0] || (unsigned)(i[1] + o1) >= (unsigned)L.n[1] ||
        (unsigned)(i[2] + o2) >= (unsigned)L.n[2])
        return -1;
    return p + o0 + (long)L.n[0] * (o1 + (long)L.n[1] * o2);
}

// ------------------------------------------------------------- thin layers
// 32 bf16 values of one pixel as four 16-byte stores / loads
__device__ inline void store_bf16x32(unsigned short* dst, const float* v)
{
#pragma unroll
    for (int g = 0; g < 4; g++) {
        bf16x8 h;
#pragma unroll
        for (int e = 0; e < 8; e++) h[e] = (short)f2bf(v[8 * g + e]);
        *reinterpret_cast<bf16x8*>(dst + 8 * g) = h;
    }
}
__device__ inline void load_bf16x32(const unsigned short* src, float* v)
{
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const bf16x8 h = *reinterpret_cast<const bf16x8*>(src + 8 * g);
#pragma unroll
        for (int e = 0; e < 8; e++) v[8 * g + e] = bf2f((unsigned short)h[e]);
    }
}

// first layer: out[p][co] = relu(b0[co] + sum_t sum_ci W0[t][co][ci] in[nb][ci]).
// TAPS = CIN = 0: extents read at run time; otherwise compile-time.  The weight
// indices are the same in every lane, so the weights come through scalar loads.
template <int TAPS, int CIN>
__global__ __launch_bounds__(256) void k_first(Lat L, int cin_rt, int taps_rt, const float* __restrict__ in,
                                               const float* __restrict__ W0, const float* __restrict__ b0,
                                               unsigned short* __restrict__ out)
{
    const int taps = TAPS ? TAPS : taps_rt, cin = CIN ? CIN : cin_rt;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < L.np; p += (long)gridDim.x * blockDim.x) {
        int ip[3];
        lat_xyz(L, p, ip);
        float acc[NN_H];
#pragma unroll
        for (int c = 0; c < NN_H; c++) acc[c] = b0[c];
        for (int t = 0; t < taps; t++) {
            const long q = lat_nb(L, p, ip, t, 1);
            if (q < 0) continue;
            for (int ci = 0; ci < cin; ci++) {
                const float v = in[q * cin + ci];
#pragma unroll
                for (int c = 0; c < NN_H; c++) acc[c] += W0[(t * NN_H + c) * cin + ci] * v;
            }
        }
#pragma unroll
        for (int c = 0; c < NN_H; c++) acc[c] = acc[c] > 0.f ? acc[c] : 0.f;
        store_bf16x32(out + p * NN_H, acc);
    }
}

// last layer: out[p] = b3 + sum_t sum_ci W3[t][ci] act[nb][ci]
template <int TAPS>
__global__ __launch_bounds__(256) void k_last(Lat L, int taps_rt, const unsigned short* __restrict__ act,
                                              const float* __restrict__ W3, const float* __restrict__ b3,
                                              float* __restrict__ out)
{
    const int taps = TAPS ? TAPS : taps_rt;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < L.np; p += (long)gridDim.x * blockDim.x) {
        int ip[3];
        lat_xyz(L, p, ip);
        float acc = b3[0];
        for (int t = 0; t < taps; t++) {
            const long q = lat_nb(L, p, ip, t, 1);
            if (q < 0) continue;
            float a[NN_H];
            load_bf16x32(act + q * NN_H, a);
#pragma unroll
            for (int c = 0; c < NN_H; c++) acc += W3[t * NN_H + c] * a[c];
        }
        out[p] = acc;
    }
}

// ------------------------------------------------------------- MFMA layer
// D[co][pixel] = sum_k W[co][k] * act[k][pixel], k = (tap, ci), 32x32 tile per
// wavefront (32 consecutive pixels), K = taps * 32 in steps of 16.
// mode 0: forward, out = relu(D + bias)            (sign +1, weights [t][co][ci])
// mode 1: input gradient, out = D * [mask > 0]     (sign -1, weights [t][ci][co])
template <int MODE, int TAPS, int NT>
__global__ __launch_bounds__(256) void k_mid(Lat L, int taps_rt, const unsigned short* __restrict__ in,
                                             const unsigned short* __restrict__ Wh, const float* __restrict__ bias,
                                             const unsigned short* __restrict__ mask, unsigned short* __restrict__ out)
{
    // The A fragments (weights) of every (tap, K half) live in LDS for the whole
    // block: wl[(t * 2 + kb) * 64 + lane] is what lane `lane` feeds the MFMA.
    // Each wavefront computes NT tiles of 32 consecutive pixels, so every A
    // fragment read from LDS feeds NT MFMAs and NT independent activation loads
    // are in flight per (tap, K half).
    extern __shared__ bf16x8 wl[];
    const int taps = TAPS ? TAPS : taps_rt;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j = lane & 31;       // pixel column of the tile / output row of W
    const int kh = lane >> 5;      // which 8-wide half of the 16-wide K step
    for (int f = threadIdx.x; f < taps * 2 * 64; f += blockDim.x) {
        const int t = f / 128, kb = (f / 64) & 1, l = f & 63;
        wl[f] = *reinterpret_cast<const bf16x8*>(Wh + ((long)t * NN_H + (l & 31)) * NN_H + kb * 16 + (l >> 5) * 8);
    }
    __syncthreads();
    const long ngroups = (L.np + 32 * NT - 1) / (32 * NT);
    for (long grp = blockIdx.x * 4L + wave; grp < ngroups; grp += (long)gridDim.x * 4) {
        long p[NT];
        int ip[NT][3];
        f32x16 acc[NT];
#pragma unroll
        for (int u = 0; u < NT; u++) {
            p[u] = (grp * NT + u) * 32 + j;
            ip[u][0] = ip[u][1] = ip[u][2] = 0;
            if (p[u] < L.np) lat_xyz(L, p[u], ip[u]);
#pragma unroll
            for (int r = 0; r < 16; r++) acc[u][r] = 0.f;
        }
        for (int t = 0; t < taps; t++) {
            long q[NT];
#pragma unroll
            for (int u = 0; u < NT; u++) q[u] = p[u] < L.np ? lat_nb(L, p[u], ip[u], t, MODE == 0 ? 1 : -1) : -1;
#pragma unroll
            for (int kb = 0; kb < 2; kb++) {
                const int c0 = kb * 16 + kh * 8;
                const bf16x8 a = wl[(t * 2 + kb) * 64 + lane];
                bf16x8 b[NT];
#pragma unroll
                for (int u = 0; u < NT; u++) {
                    if (q[u] >= 0) b[u] = *reinterpret_cast<const bf16x8*>(in + q[u] * NN_H + c0);
                    else {
#pragma unroll
                        for (int e = 0; e < 8; e++) b[u][e] = 0;
                    }
                }
#pragma unroll
                for (int u = 0; u < NT; u++) acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[u], acc[u], 0, 0, 0);
            }
        }
#pragma unroll
        for (int u = 0; u < NT; u++) {
            if (p[u] >= L.np) continue;
            // lane holds D[c][p] for c = 8*(r/4) + 4*kh + (r%4)
#pragma unroll
            for (int g = 0; g < 4; g++) {
                unsigned short v4[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int c = 8 * g + 4 * kh + e;
                    float v = acc[u][4 * g + e];
                    if (MODE == 0) { v += bias[c]; v = v > 0.f ? v : 0.f; }
                    else v = bf2f(mask[p[u] * NN_H + c]) > 0.f ? v : 0.f;
                    v4[e] = f2bf(v);
                }
                const unsigned long long pk = (unsigned long long)v4[0] | ((unsigned long long)v4[1] << 16) |
                                              ((unsigned long long)v4[2] << 32) | ((unsigned long long)v4[3] << 48);
                *reinterpret_cast<unsigned long long*>(out + p[u] * NN_H + 8 * g + 4 * kh) = pk;
            }
        }
    }
}

// two floats -> packed bf16 pair (round to nearest even, v_cvt_pk_bf16_f32)
__device__ inline unsigned int pack_bf16x2(float a, float b)
{
    bfv2 v;
    v[0] = (__bf16)a;
    v[1] = (__bf16)b;
    return __builtin_bit_cast(unsigned int, v);
}

// 32 -> 32 layer with the weights resident in registers (2D and 3D lattices).
// v_mfma_f32_16x16x32_bf16: A = 16 output channels (half h) x 32 input
// channels of one tap, B = 32 input channels x 16 pixels along axis 0.  A
// workgroup of eight waves (two per SIMD) owns a tile of 16 x 8 x 4 pixels
// (2D: 16 x 32), 32 rows of 16 pixels; wave w computes output half h = w % 2
// of the 8 rows of block w / 2 (3D: 4 x 2 rows along axes 1, 2; 2D: 8 rows
// along axis 1).  The wave reads each halo row of its block once per axis-0
// shift (3D: 6 x 4 rows x 3 shifts = 72 fragments) and multiplies it into
// every output row it reaches (up to 9 taps): 216 MFMAs per tile against 8 x
// 27 fragment reads without that reuse.  The half's 27 A fragments (27 x 16 B
// per lane) stay in registers for the whole persistent loop, one workgroup per
// CU, and the fragment reads run one group ahead of the MFMAs.  The halo tiles
// are double-buffered in LDS and filled by buffer loads into LDS (16 B per
// lane, no registers in flight): the next tile's loads are issued before the
// current tile's MFMAs, between its fragment groups.  Out-of-lattice halo pixels read
// zeros.  The tiles
// of XCD x (workgroups b with b % 8 = x: the hardware dispatches workgroups
// round-robin over the eight XCDs) are a contiguous range, so neighbouring
// tiles share their halos in one L2.
template <int D>
struct Conv32 {
    static constexpr int NW = 8;
    static constexpr int TX = 16, TY = D == 3 ? 8 : 32, TZ = D == 3 ? 4 : 1;
    static constexpr int WY = D == 3 ? 4 : 8, WZ = D == 3 ? 2 : 1;   // a wave's rows along axes 1, 2
    static constexpr int NR = WY * WZ;            // 8
    static constexpr int HX = TX + 2, HY = TY + 2, HZ = D == 3 ? TZ + 2 : 1;
    static constexpr int HP = HX * HY * HZ;       // halo pixels: 1080 (3D), 612 (2D)
    static constexpr int NG = (HP + 63) / 64;     // 64-slot groups per chunk
    static constexpr int CS = NG * 64;            // chunk stride in 16-byte slots
    static constexpr int BUF = 4 * CS;            // slots per halo buffer (3D: 69,632 bytes)
    static constexpr int TAPS = D == 3 ? 27 : 9;
    static constexpr int RZ = D == 3 ? WZ + 2 : 1, RY = WY + 2;   // halo rows of a wave's block
    static constexpr int NB = 3 * RZ * RY;        // fragment reads per tile and wave
    static constexpr int KPW = (NG + 1) / 2;      // halo loads per wave and tile (at most)
    static constexpr int FG = D == 3 ? 4 : 6;     // fragments per read group
};

// buffer_load_dwordx4 ... lds (gfx950): lane l's 16 bytes at byte offset voff of the buffer
// [base, base + nbytes) land at lds + 16 l (lds is wave-uniform); offsets at or past nbytes
// read zeros.  The host pass does not know the builtins and never calls this
__device__ inline void lds_dma16(const void* base, unsigned nbytes, unsigned voff, bf16x8* lds)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)nbytes, 0x00020000);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 16, (int)voff, 0, 0, 0);
#endif
}

// the MFMAs of one tile from halo buffer hb into acc[row] (the wave's half).  The fragments
// are read in groups of FG, one group ahead of the MFMAs that use them; stage(i) issues the
// next tile's i-th halo load of this wave between the groups
template <int MODE, int D, class Stage>
__device__ inline void conv32_tile(const bf16x8* hb, const bf16x8 (&aw)[Conv32<D>::TAPS], int wy0, int wz0,
                                   f32x4 (&acc)[Conv32<D>::NR], Stage&& stage)
{
    using T = Conv32<D>;
    constexpr int FG = T::FG, NGRP = T::NB / FG;
    static_assert(T::NB % FG == 0 && NGRP >= 2, "whole fragment groups");
    const int lane = threadIdx.x & 63;
    const bf16x8* base = hb + (lane >> 4) * T::CS + (wz0 * T::HY + wy0) * T::HX + (lane & 15);
    // fragment i: axis-0 shift dx = i / (RZ RY), halo row (rz, ry) of the wave's block
    auto slot = [](int i) {
        const int dx = i / (T::RZ * T::RY), rz = (i / T::RY) % T::RZ, ry = i % T::RY;
        return (rz * T::HY + ry) * T::HX + dx;
    };
#pragma unroll
    for (int u = 0; u < T::NR; u++) acc[u] = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 bs[2][FG];
#pragma unroll
    for (int f = 0; f < FG; f++) bs[0][f] = base[slot(f)];
#pragma unroll
    for (int gr = 0; gr < NGRP; gr++) {
        if (gr + 1 < NGRP)
#pragma unroll
            for (int f = 0; f < FG; f++) bs[(gr + 1) & 1][f] = base[slot((gr + 1) * FG + f)];
        if (gr < T::KPW) stage(gr);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int f = 0; f < FG; f++) {
            const int i = gr * FG + f;
            const int dx = i / (T::RZ * T::RY), rz = (i / T::RY) % T::RZ, ry = i % T::RY;
#pragma unroll
            for (int dz = 0; dz < (D == 3 ? 3 : 1); dz++) {
#pragma unroll
                for (int dy = 0; dy < 3; dy++) {
                    const int oz = rz - dz, oy = ry - dy;
                    if (oz >= 0 && oz < T::WZ && oy >= 0 && oy < T::WY) {
                        // forward: tap (dx, dy, dz) reads the input at offset (dx-1, dy-1, dz-1);
                        // the input gradient reads the same position with the mirrored tap's weights
                        const int t = (dz * 3 + dy) * 3 + dx, tw = MODE == 1 ? T::TAPS - 1 - t : t;
                        const int u = oz * T::WY + oy;
                        acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bfv8, aw[tw]),
                                                                         __builtin_bit_cast(bfv8, bs[gr & 1][f]),
                                                                         acc[u], 0, 0, 0);
                    }
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int i = NGRP; i < T::KPW; i++) stage(i);
}

// the destination of stores for pixels outside the lattice: every tile issues the same
// number of stores per wave, so that the wait for the next tile's halo can leave them in flight
__device__ unsigned long long kTrash[64];

template <int MODE, int D>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv32(
    Lat L, const unsigned short* __restrict__ in, const unsigned short* __restrict__ W, const float* __restrict__ bias,
    const unsigned short* __restrict__ mask, unsigned short* __restrict__ out)
{
    using T = Conv32<D>;
    // two halo buffers as separate objects, so that the loads into one are not waited for by
    // the reads of the other
    __shared__ bf16x8 buf0[T::BUF];
    __shared__ bf16x8 buf1[T::BUF];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, q = lane >> 4, h = wave & 1, blk = wave >> 1;
    const int wy0 = D == 3 ? (blk & 1) * T::WY : blk * T::WY, wz0 = D == 3 ? (blk >> 1) * T::WZ : 0;
    const int c0 = 16 * h + 4 * q;      // this lane's four output channels
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    bf16x8 aw[T::TAPS];
#pragma unroll
    for (int t = 0; t < T::TAPS; t++) aw[t] = *reinterpret_cast<const bf16x8*>(W + ((t * NN_H + 16 * h + j) * NN_H + 8 * q));
    float bv[4] = {0.f, 0.f, 0.f, 0.f};
    if (MODE == 0)
#pragma unroll
        for (int r = 0; r < 4; r++) bv[r] = bias[c0 + r];
    // halo loads: wave w fills chunk w % 4 (channels 8 (w%4) .. +7) of the halo pixels 64 k + lane
    // for the groups k = 2 i + w / 4; rel[i] is that pixel's byte offset from the halo origin,
    // hc[i] its halo coordinates (8 bits each; all ones past the halo)
    const int ch = wave & 3, kpar = wave >> 2;
    const unsigned nbytes = (unsigned)L.np * (NN_H * 2);   // below 2^31: checked at the launch
    int rel[T::KPW];
    unsigned hc[T::KPW];
#pragma unroll
    for (int i = 0; i < T::KPW; i++) {
        const int hp = 64 * (2 * i + kpar) + lane;
        const int hx = hp % T::HX, hy = (hp / T::HX) % T::HY, hz = hp / (T::HX * T::HY);
        rel[i] = (hx + nx * (hy + ny * hz)) * (NN_H * 2);
        hc[i] = hp < T::HP ? (unsigned)hx | (unsigned)hy << 8 | (unsigned)hz << 16 : 0xffffffu;
    }
    const int ntx = (nx + T::TX - 1) / T::TX, nty = (ny + T::TY - 1) / T::TY, ntz = (nz + T::TZ - 1) / T::TZ;
    const int ntile = ntx * nty * ntz;
    // issue the halo loads of the tile at (x0, y0, z0) into buf (chunk-major: 16-byte chunk c of
    // halo pixel hp at slot c * CS + hp); pixels outside the lattice or past the halo read at
    // offset nbytes (zeros)
    // halo load i of this wave for the tile at (x0, y0, z0) into buf; interior tiles need no checks
    auto stage_i = [&](int i, int x0, int y0, int z0, bf16x8* buf) {
        const int k = 2 * i + kpar;
        if (T::NG % 2 != 0 && k >= T::NG) return;   // wave-uniform
        const bool interior = x0 >= 1 && x0 + T::TX + 1 <= nx && y0 >= 1 && y0 + T::TY + 1 <= ny &&
                              (D == 2 || (z0 >= 1 && z0 + T::TZ + 1 <= nz));
        const int org = ((x0 - 1) + nx * ((y0 - 1) + ny * (D == 3 ? z0 - 1 : 0))) * (NN_H * 2) + ch * 16;
        bool ok = hc[i] != 0xffffffu;
        if (!interior) {
            const int gx = x0 - 1 + (int)(hc[i] & 0xff), gy = y0 - 1 + (int)((hc[i] >> 8) & 0xff);
            const int gz = D == 3 ? z0 - 1 + (int)(hc[i] >> 16) : 0;
            ok = ok && (unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz;
        }
        lds_dma16(in, nbytes, ok ? (unsigned)(org + rel[i]) : nbytes, buf + ch * T::CS + k * 64);
    };
    auto stage = [&](int x0, int y0, int z0, bf16x8* buf) {
#pragma unroll
        for (int i = 0; i < T::KPW; i++) stage_i(i, x0, y0, z0, buf);
    };
    int tb, tend, tstep;
    if ((gridDim.x & 7) == 0) {
        const int x = blockIdx.x & 7;
        tb = (int)((long)x * ntile / 8) + (int)(blockIdx.x >> 3);
        tend = (int)((long)(x + 1) * ntile / 8);
        tstep = gridDim.x >> 3;
    } else {
        tb = blockIdx.x;
        tend = ntile;
        tstep = gridDim.x;
    }
    if (tb < tend) stage((tb % ntx) * T::TX, (tb / ntx % nty) * T::TY, tb / (ntx * nty) * T::TZ, buf0);
    const long sy = (long)nx * NN_H, sz = (long)nx * ny * NN_H;   // row strides in elements
    for (int it = 0; tb < tend; tb += tstep, it++) {
        const bf16x8* cur = (it & 1) ? buf1 : buf0;
        bf16x8* next = (it & 1) ? buf0 : buf1;
        // this tile's halo loads (issued one iteration earlier, before the previous tile's NR
        // stores) have landed in every wave; every wave is done with the buffer the next loads fill
        if (it == 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        else asm volatile("s_waitcnt vmcnt(%0)" ::"n"(T::NR) : "memory");
        __syncthreads();
        const int tn = tb + tstep;
        const bool more = tn < tend;
        const int nx0 = (tn % ntx) * T::TX, ny0 = (tn / ntx % nty) * T::TY, nz0 = tn / (ntx * nty) * T::TZ;
        f32x4 acc[T::NR];
        // the next tile's halo loads go out between the MFMA groups
        conv32_tile<MODE, D>(cur, aw, wy0, wz0, acc, [&](int i) {
            if (more) stage_i(i, nx0, ny0, nz0, next);
        });
        const int x0 = (tb % ntx) * T::TX, y0 = (tb / ntx % nty) * T::TY, z0 = tb / (ntx * nty) * T::TZ;
        const bool xin = x0 + j < nx;
        const long p0 = (long)(x0 + j) + (long)nx * ((y0 + wy0) + (long)ny * (z0 + wz0));
#pragma unroll
        for (int u = 0; u < T::NR; u++) {
            const int oy = u % T::WY, oz = u / T::WY;
            const bool ok = xin && y0 + wy0 + oy < ny && z0 + wz0 + oz < nz;
            const long e = p0 * NN_H + oy * sy + oz * sz + c0;   // element offset of the lane's 4 channels
            float v[4];
            unsigned long long mk = 0;
            if (MODE == 1 && ok) mk = *reinterpret_cast<const unsigned long long*>(mask + e);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float x = acc[u][r];
                if (MODE == 0) {
                    x += bv[r];
                    x = x > 0.f ? x : 0.f;
                } else {
                    // bf16 bits m encode a value > 0 iff 0 < m <= 0x7f80 (sign clear, not zero, not NaN)
                    const unsigned int m = (unsigned int)(mk >> (16 * r)) & 0xffffu;
                    x = m - 1u < 0x7f80u ? x : 0.f;
                }
                v[r] = x;
            }
            uint2 pk;
            pk.x = pack_bf16x2(v[0], v[1]);
            pk.y = pack_bf16x2(v[2], v[3]);
            uint2* dst = ok ? reinterpret_cast<uint2*>(out + e) : reinterpret_cast<uint2*>(kTrash + lane);
            *dst = pk;
        }
    }
    // no LDS load may still be in flight when the workgroup's LDS is released
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// first layer (cin <= 8 input channels -> 32, 2D and 3D lattices) on
// v_mfma_f32_16x16x32_bf16 with BF16 input and weights: one K step of 32
// covers four taps x 8 channels (zero weights past cin and past the last tap),
// so 27 taps take 7 steps per output half.  The input is 16 bytes per pixel
// (8 BF16 channels).  A workgroup of four waves owns a tile of 16 x 8 x 4
// pixels (2D: 16 x 32) and wave w its 8 rows of block w (3D: 4 x 2 rows
// along axes 1, 2; 2D: rows 8w .. 8w+7), both output halves; the halo tile
// (16 bytes per pixel) is double-buffered in LDS and filled by buffer loads
// into LDS.  Lane (j, q) of a B fragment reads pixel j's channels at tap
// 4m + q, a per-lane offset into the halo tile.
template <int D>
struct ConvIn {
    static constexpr int NW = 4;
    static constexpr int TX = 16, TY = D == 3 ? 8 : 32, TZ = D == 3 ? 4 : 1;
    static constexpr int WY = D == 3 ? 4 : 8, WZ = D == 3 ? 2 : 1;
    static constexpr int NR = WY * WZ;            // 8 rows of 16 pixels per wave
    static constexpr int HX = TX + 2, HY = TY + 2, HZ = D == 3 ? TZ + 2 : 1;
    static constexpr int HP = HX * HY * HZ;
    static constexpr int NG = (HP + 63) / 64;
    static constexpr int BUF = NG * 64;           // 16-byte slots per halo buffer
    static constexpr int TAPS = D == 3 ? 27 : 9, NK = (TAPS + 3) / 4;
};

// A fragments of the first layer (K step m, half h, lane (i, q): W0[t][16 h + i][0..7] for
// tap t = 4 m + q) and this lane's halo-slot offset of tap 4 m + q (the last tap past the end)
template <int D>
__device__ inline void conv_in_setup(int cin, const float* __restrict__ W0, bf16x8 (&aw)[ConvIn<D>::NK][2],
                                     int (&toff)[ConvIn<D>::NK])
{
    using T = ConvIn<D>;
    const int lane = threadIdx.x & 63, j = lane & 15, q = lane >> 4;
#pragma unroll
    for (int m = 0; m < T::NK; m++) {
        const int t = 4 * m + q, tt = t < T::TAPS ? t : T::TAPS - 1;
        toff[m] = ((D == 3 ? tt / 9 : 0) * T::HY + (tt / 3) % 3) * T::HX + tt % 3;
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int c = 0; c < 8; c++)
                aw[m][h][c] = (short)f2bf(t < T::TAPS && c < cin ? W0[(t * NN_H + 16 * h + j) * cin + c] : 0.f);
    }
}

// the first layer's MFMAs of this wave's 8 rows (both output halves) from the channel tile cur
template <int D>
__device__ inline void conv_in_mfma(const bf16x8* cur, const bf16x8 (&aw)[ConvIn<D>::NK][2],
                                    const int (&toff)[ConvIn<D>::NK], int wy0, int wz0, f32x4 (&acc)[ConvIn<D>::NR][2])
{
    using T = ConvIn<D>;
    const int j = threadIdx.x & 15;
    const bf16x8* base = cur + (wz0 * T::HY + wy0) * T::HX + j;
    bf16x8 bf[2][T::NK];
#pragma unroll
    for (int m = 0; m < T::NK; m++) bf[0][m] = base[toff[m]];
#pragma unroll
    for (int u = 0; u < T::NR; u++) {
        if (u + 1 < T::NR) {
            const int rn = (((u + 1) / T::WY) * T::HY + (u + 1) % T::WY) * T::HX;   // the next row's halo-slot offset
#pragma unroll
            for (int m = 0; m < T::NK; m++) bf[(u + 1) & 1][m] = base[rn + toff[m]];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            acc[u][h] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < T::NK; m++)
                acc[u][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bfv8, aw[m][h]),
                                                                    __builtin_bit_cast(bfv8, bf[u & 1][m]),
                                                                    acc[u][h], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// bias, ReLU and the BF16 stores of the tile at (x0, y0, z0): 2 NR stores per wave, pixels
// outside the lattice into kTrash
template <int D>
__device__ inline void conv_in_store(const Lat& L, const f32x4 (&acc)[ConvIn<D>::NR][2], const float* __restrict__ b0,
                                     int wy0, int wz0, int x0, int y0, int z0, unsigned short* __restrict__ out)
{
    using T = ConvIn<D>;
    const int lane = threadIdx.x & 63, j = lane & 15, q = lane >> 4;
    const float4 bq0 = *reinterpret_cast<const float4*>(b0 + 4 * q), bq1 = *reinterpret_cast<const float4*>(b0 + 16 + 4 * q);
    const float bv[2][4] = {{bq0.x, bq0.y, bq0.z, bq0.w}, {bq1.x, bq1.y, bq1.z, bq1.w}};
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const long sy = (long)nx * NN_H, sz = (long)nx * ny * NN_H;
    const bool xin = x0 + j < nx;
    const long p0 = (long)(x0 + j) + (long)nx * ((y0 + wy0) + (long)ny * (z0 + wz0));
#pragma unroll
    for (int u = 0; u < T::NR; u++) {
        const int oy = u % T::WY, oz = u / T::WY;
        const bool ok = xin && y0 + wy0 + oy < ny && z0 + wz0 + oz < nz;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const long e = p0 * NN_H + oy * sy + oz * sz + 16 * h + 4 * q;
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float x = acc[u][h][r] + bv[h][r];
                v[r] = x > 0.f ? x : 0.f;
            }
            uint2 pk;
            pk.x = pack_bf16x2(v[0], v[1]);
            pk.y = pack_bf16x2(v[2], v[3]);
            uint2* dst = ok ? reinterpret_cast<uint2*>(out + e) : reinterpret_cast<uint2*>(kTrash + lane);
            *dst = pk;
        }
    }
}

// the tile range of this workgroup: tiles of XCD x (workgroups b with b % 8 = x) contiguous
__device__ inline void xcd_tiles(int ntile, int& tb, int& tend, int& tstep)
{
    if ((gridDim.x & 7) == 0) {
        const int x = blockIdx.x & 7;
        tb = (int)((long)x * ntile / 8) + (int)(blockIdx.x >> 3);
        tend = (int)((long)(x + 1) * ntile / 8);
        tstep = gridDim.x >> 3;
    } else {
        tb = blockIdx.x;
        tend = ntile;
        tstep = gridDim.x;
    }
}

template <int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv_in(
    Lat L, int cin, const unsigned short* __restrict__ in, const float* __restrict__ W0, const float* __restrict__ b0,
    unsigned short* __restrict__ out)
{
    using T = ConvIn<D>;
    __shared__ bf16x8 buf0[T::BUF];
    __shared__ bf16x8 buf1[T::BUF];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wy0 = D == 3 ? (wave & 1) * T::WY : wave * T::WY, wz0 = D == 3 ? (wave >> 1) * T::WZ : 0;
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    bf16x8 aw[T::NK][2];
    int toff[T::NK];
    conv_in_setup<D>(cin, W0, aw, toff);
    const unsigned nbytes = (unsigned)L.np * 16;   // below 2^31: checked at the launch
    const int ntx = (nx + T::TX - 1) / T::TX, nty = (ny + T::TY - 1) / T::TY, ntz = (nz + T::TZ - 1) / T::TZ;
    const int ntile = ntx * nty * ntz;
    // halo loads of the tile at (x0, y0, z0): wave w issues groups k = w, w + 4, ...; pixels
    // outside the lattice or past the halo read at offset nbytes (zeros)
    auto stage = [&](int x0, int y0, int z0, bf16x8* buf) {
#pragma unroll
        for (int i = 0; i < (T::NG + T::NW - 1) / T::NW; i++) {
            const int k = wave + T::NW * i;
            if (k >= T::NG) break;   // wave-uniform
            const int hp = 64 * k + lane;
            const int hx = hp % T::HX, hy = (hp / T::HX) % T::HY, hz = hp / (T::HX * T::HY);
            const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = D == 3 ? z0 - 1 + hz : 0;
            const bool ok = hp < T::HP && (unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny &&
                            (unsigned)gz < (unsigned)nz;
            const unsigned off = ((unsigned)gx + (unsigned)nx * ((unsigned)gy + (unsigned)ny * (unsigned)gz)) * 16u;
            lds_dma16(in, nbytes, ok ? off : nbytes, buf + k * 64);
        }
    };
    int tb, tend, tstep;
    xcd_tiles(ntile, tb, tend, tstep);
    if (tb < tend) stage((tb % ntx) * T::TX, (tb / ntx % nty) * T::TY, tb / (ntx * nty) * T::TZ, buf0);
    for (int it = 0; tb < tend; tb += tstep, it++) {
        const bf16x8* cur = (it & 1) ? buf1 : buf0;
        bf16x8* next = (it & 1) ? buf0 : buf1;
        // this tile's halo loads (issued before the previous tile's 2 NR stores) have landed in
        // every wave; every wave is done with the buffer the next loads fill
        if (it == 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        else asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * T::NR) : "memory");
        __syncthreads();
        const int tn = tb + tstep;
        if (tn < tend) stage((tn % ntx) * T::TX, (tn / ntx % nty) * T::TY, tn / (ntx * nty) * T::TZ, next);
        f32x4 acc[T::NR][2];
        conv_in_mfma<D>(cur, aw, toff, wy0, wz0, acc);
        conv_in_store<D>(L, acc, b0, wy0, wz0, (tb % ntx) * T::TX, (tb / ntx % nty) * T::TY, tb / (ntx * nty) * T::TZ,
                         out);
    }
    // no LDS load may still be in flight when the workgroup's LDS is released
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// buffer_load_dword ... lds: lane l's 4 bytes at byte offset voff of the buffer [base, base +
// nbytes) land at lds + 4 l (lds is wave-uniform); offsets at or past nbytes read zeros
__device__ inline void lds_dma4(const void* base, unsigned nbytes, unsigned voff, float* lds)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)nbytes, 0x00020000);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 4, (int)voff, 0, 0, 0);
#endif
}

// First layer of the plain learned step, from the scaled residual v alone (np FP32 values in
// lattice order) instead of a stored 8-channel input: the channels are v, v times the static
// log-coefficient channel slog (lattice order; nullptr: zero) and, for cin = 2 + d, v times
// the pixel's lattice parity along each axis, each rounded to BF16 -- bit for bit the values
// of the input k_conv_in reads.  Each wave stages the halo groups k = w, w + 4, ... of the
// next tile (v and slog, 4 bytes per pixel, by buffer loads into LDS) before this tile's
// MFMAs and converts the same groups into the next tile's 16-byte channel tile after its
// stores, so one barrier per tile orders the two channel tiles; the MFMAs and stores are
// those of k_conv_in.
template <int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv_in_v(
    Lat L, int cin, const float* __restrict__ vin, const float* __restrict__ slog, const float* __restrict__ W0,
    const float* __restrict__ b0, unsigned short* __restrict__ out)
{
    using T = ConvIn<D>;
    static_assert(T::TX % 2 == 0 && T::TY % 2 == 0 && (D == 2 || T::TZ % 2 == 0), "even tile origins");
    __shared__ bf16x8 ch0[T::BUF];
    __shared__ bf16x8 ch1[T::BUF];
    __shared__ float vb[T::BUF];
    __shared__ float lb[T::BUF];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wy0 = D == 3 ? (wave & 1) * T::WY : wave * T::WY, wz0 = D == 3 ? (wave >> 1) * T::WZ : 0;
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    bf16x8 aw[T::NK][2];
    int toff[T::NK];
    conv_in_setup<D>(cin, W0, aw, toff);
    const unsigned nbytes = (unsigned)L.np * 4;   // below 2^31: checked at the launch
    const int ntx = (nx + T::TX - 1) / T::TX, nty = (ny + T::TY - 1) / T::TY, ntz = (nz + T::TZ - 1) / T::TZ;
    const int ntile = ntx * nty * ntz;
    constexpr int KW = (T::NG + T::NW - 1) / T::NW;   // halo groups per wave (at most)
    // halo coordinates of this lane's pixel in group k = wave + NW i (8 bits each; all ones
    // past the halo)
    unsigned hc[KW];
#pragma unroll
    for (int i = 0; i < KW; i++) {
        const int hp = 64 * (wave + T::NW * i) + lane;
        const int hx = hp % T::HX, hy = (hp / T::HX) % T::HY, hz = hp / (T::HX * T::HY);
        hc[i] = hp < T::HP ? (unsigned)hx | (unsigned)hy << 8 | (unsigned)hz << 16 : 0xffffffu;
    }
    // v (and slog) of this wave's halo groups of the tile at (x0, y0, z0); pixels outside the
    // lattice or past the halo read zeros
    auto stage = [&](int x0, int y0, int z0) {
#pragma unroll
        for (int i = 0; i < KW; i++) {
            const int k = wave + T::NW * i;
            if (k >= T::NG) break;   // wave-uniform
            const int gx = x0 - 1 + (int)(hc[i] & 0xff), gy = y0 - 1 + (int)((hc[i] >> 8) & 0xff);
            const int gz = D == 3 ? z0 - 1 + (int)(hc[i] >> 16) : 0;
            const bool ok = hc[i] != 0xffffffu && (unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny &&
                            (unsigned)gz < (unsigned)nz;
            const unsigned off = ((unsigned)gx + (unsigned)nx * ((unsigned)gy + (unsigned)ny * (unsigned)gz)) * 4u;
            lds_dma4(vin, nbytes, ok ? off : nbytes, vb + 64 * k);
            if (slog) lds_dma4(slog, nbytes, ok ? off : nbytes, lb + 64 * k);
        }
    };
    // the same groups as BF16 input channels into the channel tile ch
    auto convert = [&](int x0, int y0, int z0, bf16x8* ch) {
#pragma unroll
        for (int i = 0; i < KW; i++) {
            const int k = wave + T::NW * i;
            if (k >= T::NG) break;   // wave-uniform
            const int hp = 64 * k + lane;
            const float v = vb[hp], sl = slog ? lb[hp] : 0.f;
            const float vz = v * 0.f;   // v times a zero parity channel
            // the tile origin is even on every axis, so pixel g = origin - 1 + h has parity 1
            // along an axis iff its halo coordinate h is even there
            const unsigned ev = ~hc[i];
            bf16x8 h;
            h[0] = (short)f2bf(v);
            h[1] = (short)f2bf(cin > 1 ? v * sl : 0.f);
            h[2] = (short)f2bf(cin > 2 ? ((ev & 1u) ? v : vz) : 0.f);
            h[3] = (short)f2bf(cin > 3 ? ((ev & 0x100u) ? v : vz) : 0.f);
            h[4] = (short)f2bf(D == 3 && cin > 4 ? ((ev & 0x10000u) ? v : vz) : 0.f);
            h[5] = h[6] = h[7] = 0;
            if (hp < T::HP) ch[hp] = h;
        }
    };
    int tb, tend, tstep;
    xcd_tiles(ntile, tb, tend, tstep);
    if (tb < tend) {
        const int x0 = (tb % ntx) * T::TX, y0 = (tb / ntx % nty) * T::TY, z0 = tb / (ntx * nty) * T::TZ;
        stage(x0, y0, z0);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        convert(x0, y0, z0, ch0);
    }
    for (int it = 0; tb < tend; tb += tstep, it++) {
        const bf16x8* cur = (it & 1) ? ch1 : ch0;
        bf16x8* next = (it & 1) ? ch0 : ch1;
        // every wave has converted its groups of this tile and is done with the other tile
        __syncthreads();
        const int tn = tb + tstep;
        const bool more = tn < tend;
        const int nx0 = (tn % ntx) * T::TX, ny0 = (tn / ntx % nty) * T::TY, nz0 = tn / (ntx * nty) * T::TZ;
        if (more) stage(nx0, ny0, nz0);
        f32x4 acc[T::NR][2];
        conv_in_mfma<D>(cur, aw, toff, wy0, wz0, acc);
        conv_in_store<D>(L, acc, b0, wy0, wz0, (tb % ntx) * T::TX, (tb / ntx % nty) * T::TY, tb / (ntx * nty) * T::TZ,
                         out);
        if (more) {
            // this wave's v (and slog) groups of the next tile have landed; the tile's 2 NR
            // stores, issued after them, may still be in flight
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * T::NR) : "memory");
            convert(nx0, ny0, nz0, next);
        }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// 16 bytes at byte offset voff of the buffer [base, base + nbytes), zeros at or past nbytes
__device__ inline bf16x8 load16_or_zero(const void* base, unsigned nbytes, unsigned voff)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)nbytes, 0x00020000);
    return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, 0, 0));
#else
    return bf16x8{};
#endif
}

// last layer (32 -> 1, 2D and 3D lattices) as two small products: for every halo pixel q of a
// tile, c[q][t] = W3[t] . act[q] for all taps t at once on v_mfma_f32_16x16x32_bf16 (A = 16
// halo pixels x 32 channels straight from HBM, B = 32 channels x 16 taps), into LDS; then
// out[p] = b3 + sum_t c[p + off(t)][t] for each output pixel, the taps summed in a fixed order.
// The tile is 14 pixels wide, so that its halo rows are exactly 16 pixels; halo pixels outside
// the lattice read zeros.  Four waves, two workgroups per CU.
template <int D>
struct ConvOut {
    static constexpr int NW = 4;
    static constexpr int TX = 14, TY = D == 3 ? 4 : 16, TZ = D == 3 ? 4 : 1;
    static constexpr int HX = 16, HY = TY + 2, HZ = D == 3 ? TZ + 2 : 1;
    static constexpr int NROW = HY * HZ;          // halo rows of 16 pixels: 36 (3D), 18 (2D)
    static constexpr int TAPS = D == 3 ? 27 : 9, NTB = (TAPS + 15) / 16;
    static constexpr int NOUT = TX * TY * TZ;     // 224 output pixels per tile
};

// MODE 0 writes out[p]; MODE 1 (FP32 x) and 2 (FP64 x) add the output into x at the DOF of
// pixel p instead, x[ip[p]] = x + out in FP64 rounded to x's type, where ip[p] >= 0 (the
// pixel's DOF is free; ip = -1 on Dirichlet nodes): the plain learned step without storing
// its output
template <int D, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv_out(
    Lat L, const unsigned short* __restrict__ act, const unsigned short* __restrict__ W3h,
    const float* __restrict__ b3, float* __restrict__ out, void* __restrict__ x, const int* __restrict__ ip)
{
    using T = ConvOut<D>;
    __shared__ float ct[T::NROW * 16 * T::TAPS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kq = lane >> 4;
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const unsigned nbytes = (unsigned)L.np * (NN_H * 2);   // below 2^31: checked at the launch
    // B fragments: taps 16 b + i, channels 8 kq .. 8 kq + 7 (zero past the last tap)
    bf16x8 bw[T::NTB];
#pragma unroll
    for (int b = 0; b < T::NTB; b++) {
        const int t = 16 * b + i;
        bw[b] = t < T::TAPS ? *reinterpret_cast<const bf16x8*>(W3h + t * NN_H + 8 * kq) : bf16x8{};
    }
    const float bias = b3[0];
    const int ntx = (nx + T::TX - 1) / T::TX, nty = (ny + T::TY - 1) / T::TY, ntz = (nz + T::TZ - 1) / T::TZ;
    const int ntile = ntx * nty * ntz;
    int tb, tend, tstep;
    xcd_tiles(ntile, tb, tend, tstep);
    for (; tb < tend; tb += tstep) {
        const int x0 = (tb % ntx) * T::TX, y0 = (tb / ntx % nty) * T::TY, z0 = tb / (ntx * nty) * T::TZ;
        // this thread's output pixel of phase 2; in the add modes its DOF is loaded with the halo
        // rows and x at that DOF after the MFMAs (a load issued before the halo rows would hold up
        // their wait), so that both latencies overlap the MFMAs, the barrier and the tap sums
        const int ox = threadIdx.x % T::TX, oy = (threadIdx.x / T::TX) % T::TY, oz = threadIdx.x / (T::TX * T::TY);
        const int px = x0 + ox, py = y0 + oy, pz = z0 + oz;
        const bool mine = threadIdx.x < T::NOUT && px < nx && py < ny && pz < nz;
        const long p = (long)px + (long)nx * (py + (long)ny * pz);
        int di = -1;
        if (MODE != 0 && mine) di = ip[p];
        // phase 1: wave w takes halo rows w, w + 4, ...; all of its row loads first
        constexpr int RPW = (T::NROW + T::NW - 1) / T::NW;
        bf16x8 a[RPW];
        const int gx = x0 - 1 + i;
#pragma unroll
        for (int k = 0; k < RPW; k++) {
            const int row = wave + T::NW * k;
            const int gy = y0 - 1 + row % T::HY, gz = D == 3 ? z0 - 1 + row / T::HY : 0;
            const bool ok = row < T::NROW && (unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny &&
                            (unsigned)gz < (unsigned)nz;
            const unsigned off = ((unsigned)gx + (unsigned)nx * ((unsigned)gy + (unsigned)ny * (unsigned)gz)) * (NN_H * 2) +
                                 16u * kq;
            a[k] = load16_or_zero(act, nbytes, ok ? off : nbytes);
        }
#pragma unroll
        for (int k = 0; k < RPW; k++) {
            const int row = wave + T::NW * k;
            if (row >= T::NROW) break;   // wave-uniform
#pragma unroll
            for (int b = 0; b < T::NTB; b++) {
                f32x4 d = {0.f, 0.f, 0.f, 0.f};
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bfv8, a[k]), __builtin_bit_cast(bfv8, bw[b]),
                                                            d, 0, 0, 0);
                // lane (tap column i, kq) holds c[halo pixel 4 kq + r of the row][tap 16 b + i]
                const int t = 16 * b + i;
                if (t < T::TAPS)
#pragma unroll
                    for (int r = 0; r < 4; r++) ct[(row * 16 + 4 * kq + r) * T::TAPS + t] = d[r];
            }
        }
        double xo = 0.0;
        if (MODE == 1 && di >= 0) xo = (double)static_cast<const float*>(x)[di];
        if (MODE == 2 && di >= 0) xo = static_cast<const double*>(x)[di];
        __syncthreads();
        // phase 2: one output pixel per thread
        if (mine) {
            float sum = bias;
#pragma unroll
            for (int t = 0; t < T::TAPS; t++) {
                const int dx = t % 3, dy = (t / 3) % 3, dz = D == 3 ? t / 9 : 0;
                sum += ct[(((oz + dz) * T::HY + oy + dy) * 16 + ox + dx) * T::TAPS + t];
            }
            if (MODE == 0) out[p] = sum;
            else if (MODE == 1 && di >= 0) static_cast<float*>(x)[di] = (float)(xo + (double)sum);
            else if (MODE == 2 && di >= 0) static_cast<double*>(x)[di] = xo + (double)sum;
        }
        __syncthreads();
    }
}

// BF16 copy of the input: inb[p][c] = bf16(in[p][c]) for c < cin, zeros up to 8 channels
__global__ void k_in_bf16(long np, int cin, const float* __restrict__ in, unsigned short* __restrict__ inb)
{
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
        bf16x8 v;
#pragma unroll
        for (int c = 0; c < 8; c++) v[c] = (short)f2bf(c < cin ? in[p * cin + c] : 0.f);
        *reinterpret_cast<bf16x8*>(inb + p * 8) = v;
    }
}

constexpr int MID_NT = 4;   // pixel tiles per wavefront in the 32 -> 32 layers (1D lattices)

// LDS bytes one workgroup may allocate on the current device (160 KiB on gfx950)
static long lds_per_block()
{
    static long cached = -1;
    if (cached < 0) {
        int dev = 0, v = 0;
        JFEM_HIP(hipGetDevice(&dev));
        JFEM_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        cached = v;
    }
    return cached;
}

// launch of one 32 -> 32 layer (MODE 0 forward, 1 input gradient): a persistent
// grid of at most 1024 blocks, each loading the weights into LDS once
template <int MODE>
static void launch_mid(hipStream_t st, const Lat& L, int taps, const unsigned short* in, const unsigned short* W,
                       const float* bias, const unsigned short* mask, unsigned short* out)
{
    if ((L.d == 3 || L.d == 2) && L.np * NN_H * 2 < (1L << 31) &&
        lds_per_block() >= (long)sizeof(bf16x8) * 2 * Conv32<3>::BUF) {
        // persistent: one workgroup of eight waves per CU (256 CUs), a multiple of 8 for the XCD
        // tile ranges
        const int d = L.d;
        const long tx = (L.n[0] + 15) / 16;
        const long nt = d == 3 ? tx * ((L.n[1] + 7) / 8) * ((L.n[2] + 3) / 4) : tx * ((L.n[1] + 31) / 32);
        const int grid = (int)(nt >= 256 ? 256 : (nt >= 8 ? nt & ~7L : nt));
        if (d == 3) k_conv32<MODE, 3><<<grid, Conv32<3>::NW * 64, 0, st>>>(L, in, W, bias, mask, out);
        else k_conv32<MODE, 2><<<grid, Conv32<2>::NW * 64, 0, st>>>(L, in, W, bias, mask, out);
        JFEM_HIP(hipGetLastError());
        return;
    }
    const long ngroups = (L.np + 32 * MID_NT - 1) / (32 * MID_NT);
    const int grid = (int)std::min<long>((ngroups + 3) / 4, 1024);
    const size_t lds = sizeof(bf16x8) * taps * 2 * 64;
    if (taps == 27) k_mid<MODE, 27, MID_NT><<<grid, 256, lds, st>>>(L, taps, in, W, bias, mask, out);
    else if (taps == 9) k_mid<MODE, 9, MID_NT><<<grid, 256, lds, st>>>(L, taps, in, W, bias, mask, out);
    else k_mid<MODE, 0, MID_NT><<<grid, 256, lds, st>>>(L, taps, in, W, bias, mask, out);
    JFEM_HIP(hipGetLastError());
}

// Weight gradient of a 32 -> 32 layer on MFMA:
//   dW[t][co][ci] = sum_p G[p][co] * X[nb(p,t)][ci]
// (G, X: NHWC bf16, 32 channels).  The pixels are the reduction dimension, so a
// block stages tiles of 64 pixels transposed in LDS (channel-major): G^T once
// for all its waves, and X_t^T per wave and tap.  Wave w accumulates the taps
// t = w, w+4, ...; each block covers a contiguous pixel range and writes its
// partial dW (taps x 32 x 32 floats).  The thin first and last layers use the
// same kernel on channel-padded bf16 copies of their input / output gradient.
constexpr int DW_TP = 64;   // pixels per LDS tile

template <int TAPS>
__global__ __launch_bounds__(256) void k_conv_dw(Lat L, long chunk, const unsigned short* __restrict__ G,
                                                 const unsigned short* __restrict__ X, float* __restrict__ partials)
{
    constexpr int TW = (TAPS + 3) / 4;   // taps per wave
    __shared__ __attribute__((aligned(16))) unsigned short gt[NN_H][DW_TP];
    __shared__ __attribute__((aligned(16))) unsigned short xt[4][NN_H][DW_TP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 31, kh = lane >> 5;
    const long p_begin = blockIdx.x * chunk;
    const long p_end = min(p_begin + chunk, L.np);
    f32x16 acc[TW];
#pragma unroll
    for (int tt = 0; tt < TW; tt++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[tt][r] = 0.f;
    for (long pb = p_begin; pb < p_end; pb += DW_TP) {
        // G tile: thread -> (pixel tid / 4, channels 8 * (tid % 4) .. +7), stored transposed
        {
            const int px = tid >> 2, c0 = (tid & 3) * 8;
            const long p = pb + px;
            bf16x8 v;
            if (p < p_end) v = *reinterpret_cast<const bf16x8*>(G + p * NN_H + c0);
            else {
#pragma unroll
                for (int e = 0; e < 8; e++) v[e] = 0;
            }
#pragma unroll
            for (int e = 0; e < 8; e++) gt[c0 + e][px] = (unsigned short)v[e];
        }
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < TW; tt++) {
            const int t = wave + 4 * tt;
            if (t < TAPS) {
                // X_t tile: lane -> pixel lane (64 pixels), all 32 channels, transposed
                const long p = pb + lane;
                long q = -1;
                if (p < p_end) {
                    int ip[3];
                    lat_xyz(L, p, ip);
                    q = lat_nb(L, p, ip, t, 1);
                }
#pragma unroll
                for (int c0 = 0; c0 < NN_H; c0 += 8) {
                    bf16x8 v;
                    if (q >= 0) v = *reinterpret_cast<const bf16x8*>(X + q * NN_H + c0);
                    else {
#pragma unroll
                        for (int e = 0; e < 8; e++) v[e] = 0;
                    }
#pragma unroll
                    for (int e = 0; e < 8; e++) xt[wave][c0 + e][lane] = (unsigned short)v[e];
                }
                __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0): the wave's LDS writes are done
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int ks = 0; ks < DW_TP / 16; ks++) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(&gt[i][ks * 16 + kh * 8]);
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&xt[wave][i][ks * 16 + kh * 8]);
                    acc[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[tt], 0, 0, 0);
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int tt = 0; tt < TW; tt++) {
        const int t = wave + 4 * tt;
        if (t >= TAPS) continue;
        float* out = partials + ((long)blockIdx.x * TAPS + t) * NN_H * NN_H;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int co = 8 * (r / 4) + 4 * kh + (r % 4);
            out[co * NN_H + i] = acc[tt][r];
        }
    }
}

// sum of the per-block dW partials in a fixed block order, written (add = 0) or
// added into grad[t * st + co * sco + ci] for co < nco, ci < nci
__global__ void k_sum_dw(int nblocks, int taps, const float* __restrict__ partials, float* __restrict__ grad, int nco,
                         int nci, long st, long sco, int add)
{
    const int n = taps * nco * nci;
    for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
        const int t = o / (nco * nci), co = (o / nci) % nco, ci = o % nci;
        const long src = ((long)t * NN_H + co) * NN_H + ci;
        float s = 0;
        for (int b = 0; b < nblocks; b++) s += partials[(long)b * taps * NN_H * NN_H + src];
        float* dst = grad + t * st + co * sco + ci;
        *dst = add ? *dst + s : s;
    }
}

// channel-padded bf16 copies for the thin layers' weight gradients:
// dst[p][c] = c < n ? src[p * n + c] : 0
__global__ void k_pad_bf16(long np, int n, const float* __restrict__ src, unsigned short* __restrict__ dst)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < np * NN_H; i += (long)gridDim.x * blockDim.x) {
        const long p = i / NN_H;
        const int c = (int)(i % NN_H);
        dst[i] = c < n ? f2bf(src[p * n + c]) : (unsigned short)0;
    }
}

// bias gradient partials: db[c] = sum_p g[p][c]; thread -> (pixel, 8 channels),
// 16-byte loads, per-thread sums over the chunk, then a fixed-order LDS reduction
__global__ __launch_bounds__(256) void k_bias_grad(long np, long chunk, const unsigned short* __restrict__ g,
                                                   float* __restrict__ partials)
{
    __shared__ float sh[64][NN_H + 1];
    const int c0 = (threadIdx.x & 3) * 8, px = threadIdx.x >> 2;
    const long p0 = blockIdx.x * chunk, p1 = min(p0 + chunk, np);
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long p = p0 + px; p < p1; p += 64) {
        const bf16x8 h = *reinterpret_cast<const bf16x8*>(g + p * NN_H + c0);
#pragma unroll
        for (int e = 0; e < 8; e++) acc[e] += bf2f((unsigned short)h[e]);
    }
#pragma unroll
    for (int e = 0; e < 8; e++) sh[px][c0 + e] = acc[e];
    __syncthreads();
    if (threadIdx.x < NN_H) {
        float s2 = 0;
        for (int k = 0; k < 64; k++) s2 += sh[k][threadIdx.x];
        partials[blockIdx.x * NN_H + threadIdx.x] = s2;
    }
}

// last layer backward: g2[p][ci] = [act2[p][ci] > 0] * sum_t W3[t][ci] gout[nb(p,t,-1)]
__global__ __launch_bounds__(256) void k_last_dx(Lat L, int taps, const float* __restrict__ gout,
                                                 const float* __restrict__ W3, const unsigned short* __restrict__ act2,
                                                 unsigned short* __restrict__ g2)
{
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < L.np; p += (long)gridDim.x * blockDim.x) {
        int ip[3];
        lat_xyz(L, p, ip);
        float acc[NN_H];
#pragma unroll
        for (int c = 0; c < NN_H; c++) acc[c] = 0;
        for (int t = 0; t < taps; t++) {
            const long q = lat_nb(L, p, ip, t, -1);
            if (q < 0) continue;
            const float g = gout[q];
#pragma unroll
            for (int c = 0; c < NN_H; c++) acc[c] += W3[t * NN_H + c] * g;
        }
        float a[NN_H];
        load_bf16x32(act2 + p * NN_H, a);
#pragma unroll
        for (int c = 0; c < NN_H; c++) acc[c] = a[c] > 0.f ? acc[c] : 0.f;
        store_bf16x32(g2 + p * NN_H, acc);
    }
}

// partials of sum_p g[p] over 4096-pixel chunks (one per block)
__global__ __launch_bounds__(256) void k_sum_chunks(long np, long chunk, const float* __restrict__ g,
                                                    float* __restrict__ partials)
{
    __shared__ float sh[4];
    const long p0 = blockIdx.x * chunk, p1 = min(p0 + chunk, np);
    float s = 0;
    for (long p = p0 + threadIdx.x; p < p1; p += 256) s += g[p];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// out[o] (+)= sum_b partials[b][o]: one 256-thread block per output, strided
// per-thread sums then a fixed-order tree (deterministic for a given nblocks)
__global__ __launch_bounds__(256) void k_sum_blocks(int nblocks, int nout, const float* __restrict__ partials,
                                                    float* __restrict__ out, int add)
{
    __shared__ float sh[4];
    const int o = blockIdx.x;
    float s = 0;
    for (int b = threadIdx.x; b < nblocks; b += 256) s += partials[(long)b * nout + o];
    for (int k = 32; k > 0; k >>= 1) s += __shfl_xor(s, k, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
        out[o] = add ? out[o] + r : r;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights
__global__ void k_adam(long n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2, float clip)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float gi = g[i] * clip;
        const float mi = b1 * m[i] + (1 - b1) * gi;
        const float vi = b2 * v[i] + (1 - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// bf16 copies of a 32->32 layer: Wh[t][co][ci] and the transposed, flipped
// WT[t][ci][co] used by the input gradient (tap t of the transpose reads nb(p,t,-1))
__global__ void k_refresh_bf16(int taps, const float* __restrict__ W, unsigned short* __restrict__ Wh,
                               unsigned short* __restrict__ WT)
{
    const int n = taps * NN_H * NN_H;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int t = i / (NN_H * NN_H), co = (i / NN_H) % NN_H, ci = i % NN_H;
        const unsigned short h = f2bf(W[i]);
        Wh[i] = h;
        WT[((long)t * NN_H + ci) * NN_H + co] = h;
    }
}

__global__ void k_to_bf16(int n, const float* __restrict__ src, unsigned short* __restrict__ dst)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = f2bf(src[i]);
}

// ------------------------------------------------------------- host side
static long param_count(int taps, int cin)
{
    return (long)taps * NN_H * cin + NN_H + 2L * ((long)taps * NN_H * NN_H + NN_H) + taps * NN_H + 1;
}

void nn_offsets(const jfem_nn* nn, long* o)
{
    // W0, b0, W1, b1, W2, b2, W3, b3
    const long t = nn->taps, c = nn->cin;
    o[0] = 0;
    o[1] = o[0] + t * NN_H * c;
    o[2] = o[1] + NN_H;
    o[3] = o[2] + t * NN_H * NN_H;
    o[4] = o[3] + NN_H;
    o[5] = o[4] + t * NN_H * NN_H;
    o[6] = o[5] + NN_H;
    o[7] = o[6] + t * NN_H;
    o[8] = o[7] + 1;
}

void nn_refresh(jfem_nn* nn)
{
    long o[9];
    nn_offsets(nn, o);
    hipStream_t st = nn->ctx->stream;
    k_refresh_bf16<<<64, 256, 0, st>>>(nn->taps, nn->w + o[2], nn->wh1, nn->wt1);
    k_refresh_bf16<<<64, 256, 0, st>>>(nn->taps, nn->w + o[4], nn->wh2, nn->wt2);
    k_to_bf16<<<4, 256, 0, st>>>(nn->taps * NN_H, nn->w + o[6], nn->wh3);
    JFEM_HIP(hipGetLastError());
}

void nn_reserve(jfem_nn* nn, long np)
{
    if (np <= nn->cap) return;
    for (void** p : {(void**)&nn->act0, (void**)&nn->act1, (void**)&nn->act2, (void**)&nn->g0, (void**)&nn->g1,
                     (void**)&nn->g2, (void**)&nn->in, (void**)&nn->inb, (void**)&nn->out, (void**)&nn->gout})
        (void)hipFree(*p);
    const size_t hb = sizeof(unsigned short) * NN_H * (size_t)np;
    JFEM_HIP(hipMalloc(&nn->act0, hb));
    JFEM_HIP(hipMalloc(&nn->act1, hb));
    JFEM_HIP(hipMalloc(&nn->act2, hb));
    JFEM_HIP(hipMalloc(&nn->g0, hb));
    JFEM_HIP(hipMalloc(&nn->g1, hb));
    JFEM_HIP(hipMalloc(&nn->g2, hb));
    JFEM_HIP(hipMalloc(&nn->in, sizeof(float) * nn->cin * (size_t)np));
    JFEM_HIP(hipMalloc(&nn->inb, sizeof(unsigned short) * 8 * (size_t)np));
    JFEM_HIP(hipMalloc(&nn->out, sizeof(float) * (size_t)np));
    JFEM_HIP(hipMalloc(&nn->gout, sizeof(float) * (size_t)np));
    nn->cap = np;
}

static int nn_grid(long np) { return grid_for(np, 256, 8192); }

// forward on nn->in (np x cin FP32, lattice order) into nn->out (np FP32)
bool nn_bf16_input(const jfem_nn* nn, const Lat& L)
{
    return (L.d == 2 || L.d == 3) && nn->cin <= 8 && L.np * NN_H * 2 < (1L << 31) &&
           lds_per_block() >= (long)sizeof(bf16x8) * 2 * Conv32<3>::BUF;
}

// persistent grids of the first and last layers: two workgroups per CU, a multiple of 8 for
// the XCD tile ranges
static int grid_in(const Lat& L)
{
    const long tx = (L.n[0] + 15) / 16;
    const long nt = L.d == 3 ? tx * ((L.n[1] + 7) / 8) * ((L.n[2] + 3) / 4) : tx * ((L.n[1] + 31) / 32);
    return (int)(nt >= 512 ? 512 : (nt >= 8 ? nt & ~7L : nt));
}
static int grid_out(const Lat& L)
{
    const long tox = (L.n[0] + 13) / 14;
    const long nto = L.d == 3 ? tox * ((L.n[1] + 3) / 4) * ((L.n[2] + 3) / 4) : tox * ((L.n[1] + 15) / 16);
    return (int)(nto >= 512 ? 512 : (nto >= 8 ? nto & ~7L : nto));
}

void nn_forward_plain(jfem_nn* nn, const Lat& L, const float* slog, const int* ip, void* x, bool x_double)
{
    if (!nn_bf16_input(nn, L))
        throw Error{JFEM_ERR_STATE, "nn_forward_plain: lattice without the LDS-tiled layers"};
    if (L.d == 3 && nn->cin != 2 && nn->cin != 5) throw Error{JFEM_ERR_STATE, "nn_forward_plain: 2 or 2 + d input channels"};
    if (L.d == 2 && nn->cin != 2 && nn->cin != 4) throw Error{JFEM_ERR_STATE, "nn_forward_plain: 2 or 2 + d input channels"};
    long o[9];
    nn_offsets(nn, o);
    hipStream_t st = nn->ctx->stream;
    const float *W0 = nn->w + o[0], *b0 = nn->w + o[1];
    const int gi = grid_in(L), go = grid_out(L);
    if (L.d == 3) k_conv_in_v<3><<<gi, ConvIn<3>::NW * 64, 0, st>>>(L, nn->cin, nn->in, slog, W0, b0, nn->act0);
    else k_conv_in_v<2><<<gi, ConvIn<2>::NW * 64, 0, st>>>(L, nn->cin, nn->in, slog, W0, b0, nn->act0);
    launch_mid<0>(st, L, nn->taps, nn->act0, nn->wh1, nn->w + o[3], nullptr, nn->act1);
    launch_mid<0>(st, L, nn->taps, nn->act1, nn->wh2, nn->w + o[5], nullptr, nn->act2);
    const float* b3 = nn->w + o[7];
    if (L.d == 3 && x_double) k_conv_out<3, 2><<<go, ConvOut<3>::NW * 64, 0, st>>>(L, nn->act2, nn->wh3, b3, nullptr, x, ip);
    else if (L.d == 3) k_conv_out<3, 1><<<go, ConvOut<3>::NW * 64, 0, st>>>(L, nn->act2, nn->wh3, b3, nullptr, x, ip);
    else if (x_double) k_conv_out<2, 2><<<go, ConvOut<2>::NW * 64, 0, st>>>(L, nn->act2, nn->wh3, b3, nullptr, x, ip);
    else k_conv_out<2, 1><<<go, ConvOut<2>::NW * 64, 0, st>>>(L, nn->act2, nn->wh3, b3, nullptr, x, ip);
    JFEM_HIP(hipGetLastError());
}

void nn_forward(jfem_nn* nn, const Lat& L, bool inb_ready)
{
    long o[9];
    nn_offsets(nn, o);
    hipStream_t st = nn->ctx->stream;
    const int g = nn_grid(L.np);
    const float *W0 = nn->w + o[0], *b0 = nn->w + o[1];
    if (inb_ready && !nn_bf16_input(nn, L))
        throw Error{JFEM_ERR_STATE, "nn_forward: BF16 input on a lattice without the LDS-tiled layers"};
    if (nn_bf16_input(nn, L)) {
        // 2D / 3D lattices: every layer on MFMA, the first from the BF16 input
        if (!inb_ready) k_in_bf16<<<g, 256, 0, st>>>(L.np, nn->cin, nn->in, nn->inb);
        const int gi = grid_in(L), go = grid_out(L);
        if (L.d == 3) k_conv_in<3><<<gi, ConvIn<3>::NW * 64, 0, st>>>(L, nn->cin, nn->inb, W0, b0, nn->act0);
        else k_conv_in<2><<<gi, ConvIn<2>::NW * 64, 0, st>>>(L, nn->cin, nn->inb, W0, b0, nn->act0);
        launch_mid<0>(st, L, nn->taps, nn->act0, nn->wh1, nn->w + o[3], nullptr, nn->act1);
        launch_mid<0>(st, L, nn->taps, nn->act1, nn->wh2, nn->w + o[5], nullptr, nn->act2);
        if (L.d == 3) k_conv_out<3, 0><<<go, ConvOut<3>::NW * 64, 0, st>>>(L, nn->act2, nn->wh3, nn->w + o[7], nn->out, nullptr, nullptr);
        else k_conv_out<2, 0><<<go, ConvOut<2>::NW * 64, 0, st>>>(L, nn->act2, nn->wh3, nn->w + o[7], nn->out, nullptr, nullptr);
        JFEM_HIP(hipGetLastError());
        return;
    }
    const int tc = nn->taps * 10 + nn->cin;
    switch (tc) {
    case 91: k_first<9, 1><<<g, 256, 0, st>>>(L, 1, 9, nn->in, W0, b0, nn->act0); break;
    case 92: k_first<9, 2><<<g, 256, 0, st>>>(L, 2, 9, nn->in, W0, b0, nn->act0); break;
    case 93: k_first<9, 3><<<g, 256, 0, st>>>(L, 3, 9, nn->in, W0, b0, nn->act0); break;
    case 94: k_first<9, 4><<<g, 256, 0, st>>>(L, 4, 9, nn->in, W0, b0, nn->act0); break;
    case 271: k_first<27, 1><<<g, 256, 0, st>>>(L, 1, 27, nn->in, W0, b0, nn->act0); break;
    case 272: k_first<27, 2><<<g, 256, 0, st>>>(L, 2, 27, nn->in, W0, b0, nn->act0); break;
    case 274: k_first<27, 4><<<g, 256, 0, st>>>(L, 4, 27, nn->in, W0, b0, nn->act0); break;
    case 275: k_first<27, 5><<<g, 256, 0, st>>>(L, 5, 27, nn->in, W0, b0, nn->act0); break;
    default: k_first<0, 0><<<g, 256, 0, st>>>(L, nn->cin, nn->taps, nn->in, W0, b0, nn->act0); break;
    }
    launch_mid<0>(st, L, nn->taps, nn->act0, nn->wh1, nn->w + o[3], nullptr, nn->act1);
    launch_mid<0>(st, L, nn->taps, nn->act1, nn->wh2, nn->w + o[5], nullptr, nn->act2);
    if (nn->taps == 27) k_last<27><<<g, 256, 0, st>>>(L, 27, nn->act2, nn->w + o[6], nn->w + o[7], nn->out);
    else if (nn->taps == 9) k_last<9><<<g, 256, 0, st>>>(L, 9, nn->act2, nn->w + o[6], nn->w + o[7], nn->out);
    else k_last<0><<<g, 256, 0, st>>>(L, nn->taps, nn->act2, nn->w + o[6], nn->w + o[7], nn->out);
    JFEM_HIP(hipGetLastError());
}

// backward from nn->gout (dLoss/dout) after nn_forward on the same input;
// writes (add = 0) or accumulates (add = 1) the gradient into nn->grad
void nn_backward(jfem_nn* nn, const Lat& L, int add)
{
    long o[9];
    nn_offsets(nn, o);
    hipStream_t st = nn->ctx->stream;
    const int taps = nn->taps;
    // weight gradients: nbw blocks over contiguous pixel ranges (multiples of the LDS tile)
    const long tiles = (L.np + DW_TP - 1) / DW_TP;
    const int nbw = (int)std::max<long>(1, std::min<long>(tiles / 8, 512));   // >= 8 tiles per block
    const long wchunk = (tiles + nbw - 1) / nbw * DW_TP;
    // bias gradients: 4096-pixel chunks
    const long bchunk = 4096;
    const int nbb = (int)((L.np + bchunk - 1) / bchunk);
    const long need = std::max<long>((long)nbw * taps * NN_H * NN_H, (long)nbb * NN_H);
    if (need > nn->part_cap) {
        (void)hipFree(nn->part);
        JFEM_HIP(hipMalloc(&nn->part, sizeof(float) * need));
        nn->part_cap = need;
    }
    auto conv_dw = [&](const unsigned short* G, const unsigned short* X) {
        if (taps == 27) k_conv_dw<27><<<nbw, 256, 0, st>>>(L, wchunk, G, X, nn->part);
        else if (taps == 9) k_conv_dw<9><<<nbw, 256, 0, st>>>(L, wchunk, G, X, nn->part);
        else k_conv_dw<3><<<nbw, 256, 0, st>>>(L, wchunk, G, X, nn->part);
        JFEM_HIP(hipGetLastError());
    };
    const int padg = grid_for(L.np * NN_H, 256, 16384);
    // layer 3 (last): dW3[t][ci] from the output gradient padded to 32 channels (in g0's buffer)
    k_pad_bf16<<<padg, 256, 0, st>>>(L.np, 1, nn->gout, nn->g0);
    conv_dw(nn->g0, nn->act2);
    k_sum_dw<<<grid_for((long)taps * NN_H, 64, 4096), 64, 0, st>>>(nbw, taps, nn->part, nn->grad + o[6], 1, NN_H,
                                                                  NN_H, 0, add);
    k_sum_chunks<<<nbb, 256, 0, st>>>(L.np, bchunk, nn->gout, nn->part);
    k_sum_blocks<<<1, 256, 0, st>>>(nbb, 1, nn->part, nn->grad + o[7], add);
    k_last_dx<<<nn_grid(L.np), 256, 0, st>>>(L, taps, nn->gout, nn->w + o[6], nn->act2, nn->g2);
    // layer 2: dW2 from g2 and act1, bias, input gradient masked by act1
    conv_dw(nn->g2, nn->act1);
    k_sum_dw<<<grid_for((long)taps * NN_H * NN_H, 64, 4096), 64, 0, st>>>(nbw, taps, nn->part, nn->grad + o[4], NN_H,
                                                                         NN_H, NN_H * NN_H, NN_H, add);
    k_bias_grad<<<nbb, 256, 0, st>>>(L.np, bchunk, nn->g2, nn->part);
    k_sum_blocks<<<NN_H, 256, 0, st>>>(nbb, NN_H, nn->part, nn->grad + o[5], add);
    launch_mid<1>(st, L, taps, nn->g2, nn->wt2, nullptr, nn->act1, nn->g1);
    // layer 1
    conv_dw(nn->g1, nn->act0);
    k_sum_dw<<<grid_for((long)taps * NN_H * NN_H, 64, 4096), 64, 0, st>>>(nbw, taps, nn->part, nn->grad + o[2], NN_H,
                                                                         NN_H, NN_H * NN_H, NN_H, add);
    k_bias_grad<<<nbb, 256, 0, st>>>(L.np, bchunk, nn->g1, nn->part);
    k_sum_blocks<<<NN_H, 256, 0, st>>>(nbb, NN_H, nn->part, nn->grad + o[3], add);
    launch_mid<1>(st, L, taps, nn->g1, nn->wt1, nullptr, nn->act0, nn->g0);
    // layer 0 (first): dW0[t][co][ci] from g0 and the input padded to 32 channels (in g2's buffer)
    k_pad_bf16<<<padg, 256, 0, st>>>(L.np, nn->cin, nn->in, nn->g2);
    conv_dw(nn->g0, nn->g2);
    k_sum_dw<<<grid_for((long)taps * NN_H * nn->cin, 64, 4096), 64, 0, st>>>(nbw, taps, nn->part, nn->grad + o[0],
                                                                            NN_H, nn->cin, (long)NN_H * nn->cin,
                                                                            nn->cin, add);
    k_bias_grad<<<nbb, 256, 0, st>>>(L.np, bchunk, nn->g0, nn->part);
    k_sum_blocks<<<NN_H, 256, 0, st>>>(nbb, NN_H, nn->part, nn->grad + o[1], add);
    JFEM_HIP(hipGetLastError());
}

__global__ void k_zero_range(float* __restrict__ w, long a, long b)
{
    for (long i = a + blockIdx.x * (long)blockDim.x + threadIdx.x; i < b; i += (long)gridDim.x * blockDim.x) w[i] = 0.f;
}

static void zero_biases(jfem_nn* nn)
{
    long o[9];
    nn_offsets(nn, o);
    hipStream_t st = nn->ctx->stream;
    k_zero_range<<<1, 64, 0, st>>>(nn->w, o[1], o[2]);
    k_zero_range<<<1, 64, 0, st>>>(nn->w, o[3], o[4]);
    k_zero_range<<<1, 64, 0, st>>>(nn->w, o[5], o[6]);
    k_zero_range<<<1, 64, 0, st>>>(nn->w, o[7], o[8]);
    JFEM_HIP(hipGetLastError());
}

void nn_set_bias_free(jfem_nn* nn, int on)
{
    nn->bias_free = on ? 1 : 0;
    if (on) zero_biases(nn);
}

void nn_adam_step(jfem_nn* nn, float lr, float grad_scale)
{
    nn->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)nn->step), c2 = 1.f - std::pow(b2, (float)nn->step);
    k_adam<<<grid_for(nn->nparams, 256, 1024), 256, 0, nn->ctx->stream>>>(nn->nparams, nn->w, nn->grad, nn->m, nn->v,
                                                                         lr, b1, b2, 1e-8f, c1, c2, grad_scale);
    JFEM_HIP(hipGetLastError());
    if (nn->bias_free) zero_biases(nn);
    nn_refresh(nn);
}

jfem_nn* nn_create(jfem_ctx* c, int dim, int cin, unsigned long long seed)
{
    JFEM_CHECK(dim >= 1 && dim <= 3, JFEM_ERR_ARG, "network dimension must be 1..3");
    JFEM_CHECK(cin >= 1 && cin <= 8, JFEM_ERR_ARG, "1..8 input channels");
    auto* nn = new jfem_nn();
    nn->ctx = c;
    nn->dim = dim;
    nn->cin = cin;
    nn->taps = dim == 1 ? 3 : (dim == 2 ? 9 : 27);
    nn->nparams = param_count(nn->taps, cin);
    const size_t bytes = sizeof(float) * nn->nparams;
    JFEM_HIP(hipMalloc(&nn->w, bytes));
    JFEM_HIP(hipMalloc(&nn->grad, bytes));
    JFEM_HIP(hipMalloc(&nn->m, bytes));
    JFEM_HIP(hipMalloc(&nn->v, bytes));
    const size_t hb = sizeof(unsigned short) * nn->taps * NN_H * NN_H;
    JFEM_HIP(hipMalloc(&nn->wh1, hb));
    JFEM_HIP(hipMalloc(&nn->wt1, hb));
    JFEM_HIP(hipMalloc(&nn->wh2, hb));
    JFEM_HIP(hipMalloc(&nn->wt2, hb));
    JFEM_HIP(hipMalloc(&nn->wh3, sizeof(unsigned short) * nn->taps * NN_H));
    // He initialisation from SplitMix64 + Box-Muller; biases 0, last layer scaled down
    std::vector<float> w(nn->nparams, 0.f);
    long o[9];
    nn_offsets(nn, o);
    unsigned long long s = seed;
    auto next = [&]() {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {
        double u1 = ((next() >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    };
    auto fill = [&](long a, long b, double fan_in, double gain) {
        const double sd = gain * std::sqrt(2.0 / fan_in);
        for (long i = a; i < b; i++) w[i] = (float)(sd * normal());
    };
    fill(o[0], o[1], nn->taps * cin, 1.0);
    fill(o[2], o[3], nn->taps * NN_H, 1.0);
    fill(o[4], o[5], nn->taps * NN_H, 1.0);
    fill(o[6], o[7], nn->taps * NN_H, 0.1);
    JFEM_HIP(hipMemcpy(nn->w, w.data(), bytes, hipMemcpyHostToDevice));
    JFEM_HIP(hipMemset(nn->grad, 0, bytes));
    JFEM_HIP(hipMemset(nn->m, 0, bytes));
    JFEM_HIP(hipMemset(nn->v, 0, bytes));
    nn_refresh(nn);
    JFEM_HIP(hipStreamSynchronize(c->stream));
    return nn;
}

void nn_free(jfem_nn* nn)
{
    if (!nn) return;
    (void)hipStreamSynchronize(nn->ctx->stream);
    for (void* p : {(void*)nn->w, (void*)nn->grad, (void*)nn->m, (void*)nn->v, (void*)nn->wh1, (void*)nn->wt1,
                    (void*)nn->wh2, (void*)nn->wt2, (void*)nn->wh3, (void*)nn->act0, (void*)nn->act1, (void*)nn->act2,
                    (void*)nn->g0, (void*)nn->g1, (void*)nn->g2, (void*)nn->in, (void*)nn->inb, (void*)nn->out,
                    (void*)nn->gout, (void*)nn->part})
        (void)hipFree(p);
    delete nn;
}

Lat make_lat(int d, const int* n)
{
    Lat L;
    L.d = d;
    L.np = 1;
    for (int a = 0; a < 3; a++) {
        L.n[a] = a < d ? n[a] : 1;
        JFEM_CHECK(L.n[a] >= 1, JFEM_ERR_ARG, "lattice extents must be positive");
        L.np *= L.n[a];
    }
    JFEM_CHECK(L.np < (1l << 31), JFEM_ERR_ARG, "lattice too large (2^31 nodes at most)");
    return L;
}

}  // namespace jfem

using namespace jfem;

static const char NN_MAGIC[8] = {'J', 'F', 'E', 'M', 'N', 'N', '0', '1'};

extern "C" {

int jfem_nn_create(jfem_ctx* ctx, int dim, int cin, unsigned long long seed, jfem_nn** out)
{
    API_BEGIN
    JFEM_CHECK(ctx && out, JFEM_ERR_ARG, "null argument");
    *out = nn_create(ctx, dim, cin, seed);
    API_END
}

int jfem_nn_destroy(jfem_nn* nn)
{
    API_BEGIN
    nn_free(nn);
    API_END
}

long jfem_nn_num_params(const jfem_nn* nn) { return nn ? nn->nparams : -1; }

int jfem_nn_get_params(jfem_nn* nn, float* out)
{
    API_BEGIN
    JFEM_CHECK(nn && out, JFEM_ERR_ARG, "null argument");
    JFEM_HIP(hipStreamSynchronize(nn->ctx->stream));   // work on the non-blocking context stream
    JFEM_HIP(hipMemcpy(out, nn->w, sizeof(float) * nn->nparams, hipMemcpyDeviceToHost));
    API_END
}

int jfem_nn_set_params(jfem_nn* nn, const float* in)
{
    API_BEGIN
    JFEM_CHECK(nn && in, JFEM_ERR_ARG, "null argument");
    JFEM_HIP(hipStreamSynchronize(nn->ctx->stream));   // work on the non-blocking context stream
    JFEM_HIP(hipMemcpy(nn->w, in, sizeof(float) * nn->nparams, hipMemcpyHostToDevice));
    nn_refresh(nn);
    JFEM_HIP(hipStreamSynchronize(nn->ctx->stream));
    API_END
}

int jfem_nn_forward_host(jfem_nn* nn, const int* shape, const float* in, float* out)
{
    API_BEGIN
    JFEM_CHECK(nn && shape && in && out, JFEM_ERR_ARG, "null argument");
    for (int a = 0; a < nn->dim; a++) JFEM_CHECK(shape[a] >= 1, JFEM_ERR_ARG, "empty lattice");
    Lat L = make_lat(nn->dim, shape);
    nn_reserve(nn, L.np);
    JFEM_HIP(hipMemcpy(nn->in, in, sizeof(float) * nn->cin * L.np, hipMemcpyHostToDevice));
    nn_forward(nn, L);
    JFEM_HIP(hipMemcpyAsync(out, nn->out, sizeof(float) * L.np, hipMemcpyDeviceToHost, nn->ctx->stream));
    JFEM_HIP(hipStreamSynchronize(nn->ctx->stream));
    API_END
}

int jfem_nn_gradient_host(jfem_nn* nn, const int* shape, const float* in, const float* gout, float* grad)
{
    API_BEGIN
    JFEM_CHECK(nn && shape && in && gout && grad, JFEM_ERR_ARG, "null argument");
    Lat L = make_lat(nn->dim, shape);
    nn_reserve(nn, L.np);
    JFEM_HIP(hipMemcpy(nn->in, in, sizeof(float) * nn->cin * L.np, hipMemcpyHostToDevice));
    nn_forward(nn, L);
    JFEM_HIP(hipMemcpyAsync(nn->gout, gout, sizeof(float) * L.np, hipMemcpyHostToDevice, nn->ctx->stream));
    nn_backward(nn, L, 0);
    JFEM_HIP(hipMemcpyAsync(grad, nn->grad, sizeof(float) * nn->nparams, hipMemcpyDeviceToHost, nn->ctx->stream));
    JFEM_HIP(hipStreamSynchronize(nn->ctx->stream));
    API_END
}

int jfem_nn_adam_step(jfem_nn* nn, const float* grad, float lr)
{
    API_BEGIN
    JFEM_CHECK(nn && lr > 0, JFEM_ERR_ARG, "bad argument");
    if (grad) JFEM_HIP(hipMemcpy(nn->grad, grad, sizeof(float) * nn->nparams, hipMemcpyHostToDevice));
    nn_adam_step(nn, lr, 1.0f);
    JFEM_HIP(hipStreamSynchronize(nn->ctx->stream));
    API_END
}

// checkpoint: magic, dim, cin, step, nparams, then weights, Adam m, Adam v (FP32)
int jfem_nn_save(jfem_nn* nn, const char* path)
{
    API_BEGIN
    JFEM_CHECK(nn && path, JFEM_ERR_ARG, "null argument");
    JFEM_HIP(hipStreamSynchronize(nn->ctx->stream));   // work on the non-blocking context stream
    std::vector<float> w(nn->nparams), m(nn->nparams), v(nn->nparams);
    JFEM_HIP(hipMemcpy(w.data(), nn->w, sizeof(float) * nn->nparams, hipMemcpyDeviceToHost));
    JFEM_HIP(hipMemcpy(m.data(), nn->m, sizeof(float) * nn->nparams, hipMemcpyDeviceToHost));
    JFEM_HIP(hipMemcpy(v.data(), nn->v, sizeof(float) * nn->nparams, hipMemcpyDeviceToHost));
    std::ofstream f(path, std::ios::binary);
    JFEM_CHECK(f.good(), JFEM_ERR_ARG, "cannot open checkpoint for writing");
    const long hdr[4] = {nn->dim, nn->cin, nn->step, nn->nparams};
    f.write(NN_MAGIC, 8);
    f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    f.write(reinterpret_cast<const char*>(w.data()), sizeof(float) * nn->nparams);
    f.write(reinterpret_cast<const char*>(m.data()), sizeof(float) * nn->nparams);
    f.write(reinterpret_cast<const char*>(v.data()), sizeof(float) * nn->nparams);
    JFEM_CHECK(f.good(), JFEM_ERR_STATE, "checkpoint write failed");
    API_END
}

int jfem_nn_load(jfem_ctx* ctx, const char* path, jfem_nn** out)
{
    API_BEGIN
    JFEM_CHECK(ctx && path && out, JFEM_ERR_ARG, "null argument");
    std::ifstream f(path, std::ios::binary);
    JFEM_CHECK(f.good(), JFEM_ERR_ARG, "cannot open checkpoint");
    char magic[8];
    long hdr[4];
    f.read(magic, 8);
    f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    JFEM_CHECK(f.good() && std::memcmp(magic, NN_MAGIC, 8) == 0, JFEM_ERR_ARG, "not a network checkpoint");
    jfem_nn* nn = nn_create(ctx, (int)hdr[0], (int)hdr[1], 0);
    if (nn->nparams != hdr[3]) { nn_free(nn); throw Error{JFEM_ERR_ARG, "checkpoint parameter count mismatch"}; }
    std::vector<float> w(nn->nparams), m(nn->nparams), v(nn->nparams);
    f.read(reinterpret_cast<char*>(w.data()), sizeof(float) * nn->nparams);
    f.read(reinterpret_cast<char*>(m.data()), sizeof(float) * nn->nparams);
    f.read(reinterpret_cast<char*>(v.data()), sizeof(float) * nn->nparams);
    if (!f.good()) { nn_free(nn); throw Error{JFEM_ERR_ARG, "truncated checkpoint"}; }
    nn->step = (int)hdr[2];
    JFEM_HIP(hipMemcpy(nn->w, w.data(), sizeof(float) * nn->nparams, hipMemcpyHostToDevice));
    JFEM_HIP(hipMemcpy(nn->m, m.data(), sizeof(float) * nn->nparams, hipMemcpyHostToDevice));
    JFEM_HIP(hipMemcpy(nn->v, v.data(), sizeof(float) * nn->nparams, hipMemcpyHostToDevice));
    nn_refresh(nn);
    JFEM_HIP(hipStreamSynchronize(ctx->stream));
    *out = nn;
    API_END
}

}  // extern "C"
