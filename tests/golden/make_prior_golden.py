"""Writes prior_conv_golden.npz: torch.nn.functional.conv2d (CPU, float64) forward and autograd gradients of a
two-layer conv3x3 network with ReLU, and float32 -> bfloat16 casts of edge values. torch is a library here, not the
reference; the fixture pins oracle/prior_ref.py. Run: python tests/golden/make_prior_golden.py"""
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(2024)
    n, h, w, c0, c1, c2 = 2, 5, 7, 3, 4, 2
    x = rng.standard_normal((n, h, w, c0))
    w1 = rng.standard_normal((c1, 3, 3, c0)) * 0.4
    b1 = rng.standard_normal(c1) * 0.1
    w2 = rng.standard_normal((c2, 3, 3, c1)) * 0.4
    b2 = rng.standard_normal(c2) * 0.1
    t = (rng.random((n, h, w, c2)) < 0.3).astype(np.float64)
    m = (rng.random((n, h, w)) < 0.8).astype(np.float64)

    tx = torch.tensor(x).permute(0, 3, 1, 2)
    tw1 = torch.tensor(w1).permute(0, 3, 1, 2).requires_grad_()
    tb1 = torch.tensor(b1).requires_grad_()
    tw2 = torch.tensor(w2).permute(0, 3, 1, 2).requires_grad_()
    tb2 = torch.tensor(b2).requires_grad_()
    a1 = torch.relu(torch.nn.functional.conv2d(tx, tw1, tb1, padding=1))
    z = torch.nn.functional.conv2d(a1, tw2, tb2, padding=1).permute(0, 2, 3, 1)
    tm = torch.tensor(m)[..., None].expand_as(z)
    l = torch.nn.functional.binary_cross_entropy_with_logits(z, torch.tensor(t), reduction="none")
    loss = (l * tm).sum() / tm.sum()
    loss.backward()

    edge = np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 1.0078125 + 2**-17, -3.0e38, 3.4e38, 1e-40, -1e-42,
                     np.inf, -np.inf, np.nan, 65504.0, 2**-126, 0.1, -0.1], np.float32)
    edge = np.concatenate([edge, rng.standard_normal(64).astype(np.float32) * 100,
                           (np.arange(64, dtype=np.uint32) * 0x8000 + 0x3F800000).view(np.float32)])
    ebf = torch.tensor(edge).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.savez(os.path.join(HERE, "prior_conv_golden.npz"), x=x, w1=w1, b1=b1, w2=w2, b2=b2, t=t, m=m,
             z=z.detach().numpy(), loss=np.array(loss.item()),
             gw1=tw1.grad.permute(0, 2, 3, 1).numpy(), gb1=tb1.grad.numpy(),
             gw2=tw2.grad.permute(0, 2, 3, 1).numpy(), gb2=tb2.grad.numpy(), edge=edge, edge_bf16=ebf)


if __name__ == "__main__":
    main()
