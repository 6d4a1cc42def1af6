"""The learned-smoother network on the GPU (BF16 MFMA, FP32 accumulation)
against the FP64 oracle: forward and parameter gradients.  The tolerances come
from an error model: the same FP64 oracle with BF16 rounding (unit roundoff
2^-8) where the device stores BF16 (weights of the 32->32 and last layers,
activations, back-propagated gradients).  Device and model errors are sums of
many independent roundings of the same size, so the device error of every
output and gradient block stays within 3x the model's (plus 1e-3); a wrong
tap, channel or term gives errors of order one."""
import os

import numpy as np
import pytest

from oracle import nn_oracle as NO

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("d,shape,cin", [(2, (37, 29), 2), (2, (64, 64), 3), (3, (9, 7, 11), 4), (2, (5, 3), 1)])
def test_forward_and_gradient_vs_oracle(jf, d, shape, cin):
    net = jf.SmootherNet(d, cin, seed=11)
    p = net.get_params().astype(np.float64)
    np_ = int(np.prod(shape))
    rng = np.random.default_rng(5)
    inp = rng.standard_normal((np_, cin))
    out = net.forward(shape, inp)
    ref, _ = NO.forward(p, shape, inp, cin)
    model, _ = NO.forward(p, shape, inp, cin, emulate_bf16=True)
    assert rel(out, ref) <= 3 * rel(model, ref) + 1e-3, (rel(out, ref), rel(model, ref))
    gout = rng.standard_normal(np_)
    g = net.gradient(shape, inp, gout)
    gr = NO.gradient(p, shape, inp, gout, cin)
    gm = NO.gradient(p, shape, inp, gout, cin, emulate_bf16=True)
    assert rel(g, gr) <= 3 * rel(gm, gr) + 1e-3, (rel(g, gr), rel(gm, gr))
    # per block (each layer's weights and bias), so that a wrong layer cannot hide behind the others
    sizes = [3 ** d * 32 * cin, 32, 3 ** d * 1024, 32, 3 ** d * 1024, 32, 3 ** d * 32, 1]
    o = 0
    for i, s in enumerate(sizes):
        if np.linalg.norm(gr[o:o + s]) > 1e-6:
            e, em = rel(g[o:o + s], gr[o:o + s]), rel(gm[o:o + s], gr[o:o + s])
            assert e <= 3 * em + 1e-3, (i, e, em)
        o += s


@pytest.mark.parametrize("d", [2, 3])
def test_torch_fixture_on_gpu(jf, d):
    f = np.load(os.path.join(GOLDEN, f"conv_net_{d}d.npz"))
    shape, cin = tuple(int(s) for s in f["shape"]), int(f["cin"])
    net = jf.SmootherNet(d, cin)
    net.set_params(f["params"])
    p = np.asarray(f["params"], np.float64)
    model, _ = NO.forward(p, shape, f["inp"], cin, emulate_bf16=True)
    gm = NO.gradient(p, shape, f["inp"], f["gout"], cin, emulate_bf16=True)
    assert rel(net.forward(shape, f["inp"]), f["out"]) <= 3 * rel(model, f["out"]) + 1e-3
    assert rel(net.gradient(shape, f["inp"], f["gout"]), f["grad"]) <= 3 * rel(gm, f["grad"]) + 1e-3


def test_checkpoint_roundtrip_and_adam(jf, tmp_path):
    net = jf.SmootherNet(2, 2, seed=3)
    g = np.random.default_rng(0).standard_normal(net.num_params()).astype(np.float32)
    p0 = net.get_params()
    net.adam_step(g, lr=1e-3)
    p1 = net.get_params()
    # first Adam step moves every weight by lr * sign(g) (bias-corrected moments)
    assert np.allclose(p1 - p0, -1e-3 * np.sign(g), atol=1e-6)
    path = str(tmp_path / "net.ckpt")
    net.save(path)
    net2 = jf.SmootherNet.load(path)
    assert np.array_equal(net2.get_params(), p1)
    net.adam_step(g, lr=1e-3)
    net2.adam_step(g, lr=1e-3)
    assert np.array_equal(net.get_params(), net2.get_params())   # optimiser state restored
