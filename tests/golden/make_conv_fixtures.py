"""Writes tests/golden/conv_net_2d.npz and conv_net_3d.npz with torch's CPU
conv2d/conv3d autograd (float64) for the learned-smoother network: an
independent reference for oracle/nn_oracle.py.  Run in the build container:
    python tests/golden/make_conv_fixtures.py
"""
import os

import numpy as np
import torch

H = 32
HERE = os.path.dirname(os.path.abspath(__file__))


def make(d, shape, cin, seed):
    rng = np.random.default_rng(seed)
    T = 3 ** d
    conv = torch.nn.functional.conv2d if d == 2 else torch.nn.functional.conv3d
    ks = (3,) * d
    Ws = [torch.tensor(rng.standard_normal((H, cin) + ks) * 0.3, requires_grad=True),
          torch.tensor(rng.standard_normal((H, H) + ks) * 0.1, requires_grad=True),
          torch.tensor(rng.standard_normal((H, H) + ks) * 0.1, requires_grad=True),
          torch.tensor(rng.standard_normal((1, H) + ks) * 0.1, requires_grad=True)]
    bs = [torch.tensor(rng.standard_normal(H) * 0.1, requires_grad=True),
          torch.tensor(rng.standard_normal(H) * 0.1, requires_grad=True),
          torch.tensor(rng.standard_normal(H) * 0.1, requires_grad=True),
          torch.tensor(rng.standard_normal(1) * 0.1, requires_grad=True)]
    # torch layout: (batch, channel, spatial reversed: axis d-1 ... axis 0)
    inp = rng.standard_normal(tuple(reversed(shape)) + (cin,))
    x = torch.tensor(np.moveaxis(inp, -1, 0)[None])
    h = x
    for i in range(4):
        h = conv(h, Ws[i], bs[i], padding=1)
        if i < 3:
            h = torch.relu(h)
    out = h[0, 0]
    gout = rng.standard_normal(out.shape)
    (out * torch.tensor(gout)).sum().backward()

    def ours(W):
        # torch W[co][ci][k_{d-1}]...[k_0] -> ours W[t][co][ci] with t = sum_a k_a 3^a
        Wn = W.detach().numpy()
        out = np.zeros((T,) + Wn.shape[:2])
        for t in range(T):
            k = [(t // 3 ** a) % 3 for a in range(d)]
            out[t] = Wn[(slice(None), slice(None)) + tuple(reversed(k))]
        return out

    def ours_g(W):
        class G:
            pass
        g = G()
        g.detach = lambda: W.grad
        return ours(g)

    params = np.concatenate([np.concatenate([ours(Ws[i]).reshape(-1), bs[i].detach().numpy()]) for i in range(4)])
    grads = np.concatenate([np.concatenate([ours_g(Ws[i]).reshape(-1), bs[i].grad.numpy()]) for i in range(4)])
    np.savez(os.path.join(HERE, f"conv_net_{d}d.npz"), shape=np.array(shape), cin=cin, params=params,
             inp=inp.reshape(-1, cin), out=out.detach().numpy().reshape(-1), gout=gout.reshape(-1), grad=grads)


if __name__ == "__main__":
    make(2, (7, 5), 2, 1)
    make(3, (4, 3, 5), 4, 2)
    print("written")
