// Learned-smoother network: handle layout and the device entry points shared
// by the network kernels (nn.hip) and the multigrid smoother (mg.hip).
#pragma once

#include "common.hpp"

namespace jfem {

constexpr int NN_H = 32;   // hidden channels

// structured lattice of pixels (nodes), axis 0 fastest in memory
struct Lat {
    int d;
    int n[3];
    long np;
};

Lat make_lat(int d, const int* n);

typedef short bf16x8 __attribute__((ext_vector_type(8)));

// round-to-nearest-even float -> bf16 bits
__host__ __device__ inline unsigned short f2bf(float f)
{
    unsigned int u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);   // inf/nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short h)
{
    unsigned int u = (unsigned int)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

}  // namespace jfem

struct jfem_nn {
    jfem_ctx* ctx = nullptr;
    int dim = 2, cin = 2, taps = 9;
    long nparams = 0;
    int step = 0;                       // Adam step counter
    int bias_free = 0;                  // biases held at zero: the network is positively homogeneous
    float *w = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;   // FP32 master weights, gradient, Adam moments
    unsigned short *wh1 = nullptr, *wt1 = nullptr, *wh2 = nullptr, *wt2 = nullptr;   // bf16 copies of the 32->32 layers
    unsigned short* wh3 = nullptr;      // bf16 copy of the last layer's weights
    // work buffers for np <= cap pixels
    long cap = 0;
    unsigned short *act0 = nullptr, *act1 = nullptr, *act2 = nullptr;   // hidden activations (post-ReLU), NHWC bf16
    unsigned short *g0 = nullptr, *g1 = nullptr, *g2 = nullptr;         // gradients w.r.t. hidden pre-activations
    float* in = nullptr;     // np x cin input (lattice order; v of nn_forward_plain in the first np values)
    unsigned short* inb = nullptr;   // np x 8 BF16 copy of the input (zeros past cin)
    float* out = nullptr;    // np output
    float* gout = nullptr;   // dLoss/dout
    float* part = nullptr;   // per-block weight-gradient partials
    long part_cap = 0;
};

namespace jfem {
void nn_reserve(jfem_nn* nn, long np);
// forward on nn->in, or, when inb_ready, on nn->inb already written by the
// caller (only valid where nn_bf16_input(nn, L) holds)
void nn_forward(jfem_nn* nn, const Lat& L, bool inb_ready = false);
// true where the forward reads its input as the BF16 copy nn->inb
bool nn_bf16_input(const jfem_nn* nn, const Lat& L);
// the plain learned step x += N(v) on a lattice where nn_bf16_input holds, for a network with
// 2 or 2 + d input channels: v (np FP32 values, lattice order) in the first np values of
// nn->in, the input channels (v, v slog, v times the lattice parity along each axis) formed
// while the first layer stages its halo (slog: lattice order, nullptr for zeros), and the
// output added into x (FP32, or FP64 where x_double) at the DOF ip[p] of every pixel p with
// ip[p] >= 0
void nn_forward_plain(jfem_nn* nn, const Lat& L, const float* slog, const int* ip, void* x, bool x_double);
void nn_backward(jfem_nn* nn, const Lat& L, int add);
void nn_adam_step(jfem_nn* nn, float lr, float grad_scale);
void nn_set_bias_free(jfem_nn* nn, int on);
}  // namespace jfem
