"""FP64 NumPy restatement of the learned-smoother network (test infrastructure only).

Four 3^d-tap convolutions on a lattice (axis 0 fastest in memory), channels
cin -> 32 -> 32 -> 32 -> 1 with ReLU between, zero padding; parameters in
the C ABI's order W0[t][32][cin], b0, W1[t][32][32], b1, W2, b2, W3[t][32], b3.
Pinned by finite differences (tests/test_nn_oracle.py) and by fixtures that
torch's CPU conv2d autograd wrote (tests/golden/make_conv_fixtures.py).
"""
from __future__ import annotations

import itertools

import numpy as np

H = 32


def taps_of(d):
    return 3 ** d


def offsets(d):
    """tap t -> offset per lattice axis (axis 0 fastest in t)"""
    out = []
    for t in range(3 ** d):
        out.append(tuple(((t // 3 ** a) % 3) - 1 for a in range(d)))
    return out


def split(params, d, cin):
    T = taps_of(d)
    sizes = [T * H * cin, H, T * H * H, H, T * H * H, H, T * H, 1]
    shapes = [(T, H, cin), (H,), (T, H, H), (H,), (T, H, H), (H,), (T, 1, H), (1,)]
    out, o = [], 0
    for s, sh in zip(sizes, shapes):
        out.append(np.asarray(params[o:o + s], np.float64).reshape(sh))
        o += s
    assert o == len(params)
    return out


def nparams(d, cin):
    T = taps_of(d)
    return T * H * cin + H + 2 * (T * H * H + H) + T * H + 1


def _shift(x, off):
    """y[p] = x[p + off] with zero padding; x has lattice axes reversed (numpy C order) + channel axis"""
    d = len(off)
    y = np.zeros_like(x)
    src = [slice(None)] * x.ndim
    dst = [slice(None)] * x.ndim
    for a in range(d):
        ax = d - 1 - a           # numpy axis of lattice axis a
        o = off[a]
        n = x.shape[ax]
        if o > 0:
            src[ax] = slice(o, n)
            dst[ax] = slice(0, n - o)
        elif o < 0:
            src[ax] = slice(0, n + o)
            dst[ax] = slice(-o, n)
        else:
            src[ax] = slice(0, n)
            dst[ax] = slice(0, n)
    y[tuple(dst)] = x[tuple(src)]
    return y


def conv(x, W, b, d):
    """x: lattice (reversed axes) x cin; W: (T, cout, cin)"""
    offs = offsets(d)
    out = np.zeros(x.shape[:-1] + (W.shape[1],))
    for t, off in enumerate(offs):
        out += _shift(x, off) @ W[t].T
    if b is not None:
        out += b
    return out


def conv_transpose(g, W, d):
    """input gradient of conv: g (..., cout) -> (..., cin)"""
    offs = offsets(d)
    out = np.zeros(g.shape[:-1] + (W.shape[2],))
    for t, off in enumerate(offs):
        out += _shift(g, tuple(-o for o in off)) @ W[t]
    return out


def conv_wgrad(g, x, d):
    """dW[t][co][ci] = sum_p g[p][co] x[p + off(t)][ci]"""
    offs = offsets(d)
    gf = g.reshape(-1, g.shape[-1])
    return np.stack([gf.T @ _shift(x, off).reshape(-1, x.shape[-1]) for off in offs])


def bf16(a):
    """round to bfloat16 (8 significant bits, round to nearest, ties to even; unit roundoff 2^-8),
    returned as float64: the error model of the device's BF16 storage"""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(a))


def forward(params, shape, inp, cin, emulate_bf16=False):
    """FP64 forward; emulate_bf16 rounds to BF16 where the device stores BF16: the input, the weights
    of every layer and the activations after each ReLU (biases and accumulation stay exact)"""
    d = len(shape)
    W0, b0, W1, b1, W2, b2, W3, b3 = split(params, d, cin)
    r = bf16 if emulate_bf16 else (lambda v: v)
    W0, W1, W2, W3 = r(W0), r(W1), r(W2), r(W3)
    x = r(np.asarray(inp, np.float64).reshape(tuple(reversed(shape)) + (cin,)))
    z0 = conv(x, W0, b0, d)
    a0 = r(np.maximum(z0, 0))
    z1 = conv(a0, W1, b1, d)
    a1 = r(np.maximum(z1, 0))
    z2 = conv(a1, W2, b2, d)
    a2 = r(np.maximum(z2, 0))
    out = conv(a2, W3, b3, d)[..., 0]
    cache = dict(x=x, a0=a0, a1=a1, a2=a2, W=(W0, W1, W2, W3), d=d)
    return out.reshape(-1), cache


def gradient(params, shape, inp, gout, cin, emulate_bf16=False):
    """gradient of sum(gout * forward(inp)) w.r.t. the parameters (flat, C-ABI order);
    emulate_bf16 also rounds the output gradient, the back-propagated gradients of each layer
    and the input of the first weight gradient to BF16, as the device stores them"""
    out, c = forward(params, shape, inp, cin, emulate_bf16)
    r = bf16 if emulate_bf16 else (lambda v: v)
    d = c["d"]
    W0, W1, W2, W3 = c["W"]
    g = np.asarray(gout, np.float64).reshape(c["a2"].shape[:-1] + (1,))
    dW3 = conv_wgrad(r(g), c["a2"], d)
    db3 = g.sum(axis=tuple(range(g.ndim - 1)))
    g2 = r(conv_transpose(g, W3, d) * (c["a2"] > 0))
    dW2 = conv_wgrad(g2, c["a1"], d)
    db2 = g2.sum(axis=tuple(range(g2.ndim - 1)))
    g1 = r(conv_transpose(g2, W2, d) * (c["a1"] > 0))
    dW1 = conv_wgrad(g1, c["a0"], d)
    db1 = g1.sum(axis=tuple(range(g1.ndim - 1)))
    g0 = r(conv_transpose(g1, W1, d) * (c["a0"] > 0))
    dW0 = conv_wgrad(g0, r(c["x"]), d)
    db0 = g0.sum(axis=tuple(range(g0.ndim - 1)))
    return np.concatenate([a.reshape(-1) for a in (dW0, db0, dW1, db1, dW2, db2, dW3, db3)])


def init_params(d, cin, seed=0):
    rng = np.random.default_rng(seed)
    T = taps_of(d)
    parts = [rng.standard_normal(T * H * cin) * np.sqrt(2 / (T * cin)), rng.standard_normal(H) * 0.1,
             rng.standard_normal(T * H * H) * np.sqrt(2 / (T * H)), rng.standard_normal(H) * 0.1,
             rng.standard_normal(T * H * H) * np.sqrt(2 / (T * H)), rng.standard_normal(H) * 0.1,
             rng.standard_normal(T * H) * np.sqrt(2 / (T * H)), rng.standard_normal(1) * 0.1]
    return np.concatenate(parts)
