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
