"""Pins the FP64 network oracle (oracle/nn_oracle.py): torch CPU autograd
fixtures and central finite differences.  CPU only."""
import os

import numpy as np
import pytest

from oracle import nn_oracle as NO

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("d", [2, 3])
def test_oracle_matches_torch_fixture(d):
    f = np.load(os.path.join(GOLDEN, f"conv_net_{d}d.npz"))
    shape, cin = tuple(int(s) for s in f["shape"]), int(f["cin"])
    out, _ = NO.forward(f["params"], shape, f["inp"], cin)
    assert np.max(np.abs(out - f["out"])) <= 1e-12 * np.max(np.abs(f["out"]))
    g = NO.gradient(f["params"], shape, f["inp"], f["gout"], cin)
    assert np.max(np.abs(g - f["grad"])) <= 1e-11 * np.max(np.abs(f["grad"]))


def test_oracle_gradient_finite_differences():
    d, shape, cin = 2, (6, 5), 2
    p = NO.init_params(d, cin, seed=3)
    rng = np.random.default_rng(4)
    inp = rng.standard_normal((30, cin))
    gout = rng.standard_normal(30)
    g = NO.gradient(p, shape, inp, gout, cin)
    f = lambda q: float(gout @ NO.forward(q, shape, inp, cin)[0])
    idx = rng.choice(len(p), 25, replace=False)
    for i in idx:
        h = 1e-6
        e = np.zeros_like(p)
        e[i] = h
        fd = (f(p + e) - f(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_tap_order_matters():
    """a transposed tap order (axis 1 fastest) must break the fixture"""
    f = np.load(os.path.join(GOLDEN, "conv_net_2d.npz"))
    shape, cin = tuple(int(s) for s in f["shape"]), int(f["cin"])
    p = f["params"].copy()
    T = 9
    W0 = p[:T * 32 * cin].reshape(T, 32, cin)
    perm = [(t % 3) * 3 + t // 3 for t in range(9)]
    p[:T * 32 * cin] = W0[perm].reshape(-1)
    out, _ = NO.forward(p, shape, f["inp"], cin)
    assert np.max(np.abs(out - f["out"])) > 1e-3


def test_bf16_rounding_model():
    """the BF16 rounding of the error model: 8 significant bits, nearest, ties to even, exact
    on representable values, relative error at most the unit roundoff 2^-8"""
    from oracle.nn_oracle import bf16
    assert bf16(1.0 + 2 ** -8) == 1.0                  # halfway, even neighbour below
    assert bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6    # halfway, even neighbour above
    assert bf16(1.0 + 2 ** -8 + 2 ** -12) == 1.0 + 2 ** -7
    assert bf16(-3.0) == -3.0 and bf16(0.0) == 0.0
    x = np.random.default_rng(1).standard_normal(10000) * 100
    assert np.max(np.abs(bf16(x) - x) / np.abs(x)) <= 2 ** -8
    assert np.array_equal(bf16(bf16(x)), bf16(x))


def test_bf16_emulation_changes_results_at_bf16_scale():
    """the emulated forward and gradient differ from the exact FP64 ones at the BF16 scale
    (up to 2^-8 relative per rounding, compounded over the layers): far above FP64 rounding,
    far below the order-one errors of a wrong tap, channel or term"""
    from oracle import nn_oracle as NO
    d, shape, cin = 2, (9, 7), 2
    p = NO.init_params(d, cin, 3)
    rng = np.random.default_rng(4)
    inp = rng.standard_normal((63, cin))
    gout = rng.standard_normal(63)
    o, _ = NO.forward(p, shape, inp, cin)
    oe, _ = NO.forward(p, shape, inp, cin, emulate_bf16=True)
    g = NO.gradient(p, shape, inp, gout, cin)
    ge = NO.gradient(p, shape, inp, gout, cin, emulate_bf16=True)
    r1 = np.linalg.norm(oe - o) / np.linalg.norm(o)
    r2 = np.linalg.norm(ge - g) / np.linalg.norm(g)
    assert 1e-5 < r1 < 0.2 and 1e-5 < r2 < 0.2, (r1, r2)
