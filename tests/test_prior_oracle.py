"""Pins the CPU restatement of the prior network (oracle/prior_ref.py): against torch's conv2d and autograd (a
committed fixture and live CPU calls), against central finite differences, and BF16 rounding bit by bit."""
import os

import numpy as np
import pytest
import torch

from oracle import prior_ref as R

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "prior_conv_golden.npz")


def test_bf16_rounding_matches_torch_cast_on_the_fixture():
    g = np.load(GOLDEN)
    got = R.bf16_bits(g["edge"])
    want = g["edge_bf16"]
    nan = np.isnan(g["edge"])
    assert np.array_equal(got[~nan], want[~nan])
    assert np.all((got[nan] & 0x7F80) == 0x7F80) and np.all(got[nan] & 0x7F)


def test_bf16_rounding_ties_go_to_even():
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: even mantissa (1.0) wins; 1 + 3 * 2^-8 goes up to 1 + 2^-6
    assert R.bf16_round(np.float32(1 + 2 ** -8)) == np.float32(1.0)
    assert R.bf16_round(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)
    assert R.bf16_round(np.float32(1 + 2 ** -8 + 2 ** -20)) == np.float32(1 + 2 ** -7)
    x = np.random.default_rng(1).standard_normal(10000).astype(np.float32)
    assert np.array_equal(R.bf16_round(x), torch.tensor(x).to(torch.bfloat16).float().numpy())


def _golden_params(g):
    return [(g["w1"], g["b1"]), (g["w2"], g["b2"])]


def test_forward_and_gradients_match_the_torch_fixture():
    g = np.load(GOLDEN)
    params = _golden_params(g)
    z, _ = R.forward(params, g["x"], quant=False)
    np.testing.assert_allclose(z, g["z"], rtol=1e-12, atol=1e-12)
    loss, grads = R.loss_grad(params, g["x"], g["t"], g["m"], quant=False)
    assert abs(loss - float(g["loss"])) < 1e-12
    np.testing.assert_allclose(grads[0][0], g["gw1"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(grads[0][1], g["gb1"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(grads[1][0], g["gw2"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(grads[1][1], g["gb2"], rtol=1e-10, atol=1e-12)


def test_six_layer_network_matches_live_torch():
    rng = np.random.default_rng(5)
    params = R.random_params(rng, 20, 32, 6, 4)
    x = rng.standard_normal((1, 9, 13, 20))
    t = (rng.random((1, 9, 13, 4)) < 0.2).astype(np.float64)
    m = np.ones((1, 9, 13))
    tx = torch.tensor(x).permute(0, 3, 1, 2)
    tp = [(torch.tensor(w.astype(np.float64)).permute(0, 3, 1, 2).requires_grad_(),
           torch.tensor(b.astype(np.float64)).requires_grad_()) for w, b in params]
    a = tx
    for L, (w, b) in enumerate(tp):
        a = torch.nn.functional.conv2d(a, w, b, padding=1)
        if L < len(tp) - 1:
            a = torch.relu(a)
    z = a.permute(0, 2, 3, 1)
    loss = torch.nn.functional.binary_cross_entropy_with_logits(z, torch.tensor(t))
    loss.backward()
    zz, _ = R.forward(params, x, quant=False)
    np.testing.assert_allclose(zz, z.detach().numpy(), rtol=1e-10, atol=1e-10)
    l2, grads = R.loss_grad(params, x, t, m, quant=False)
    assert abs(l2 - loss.item()) < 1e-10
    for (gw, gb), (w, b) in zip(grads, tp):
        np.testing.assert_allclose(gw, w.grad.permute(0, 2, 3, 1).numpy(), rtol=1e-8, atol=1e-12)
        np.testing.assert_allclose(gb, b.grad.numpy(), rtol=1e-8, atol=1e-12)


def test_gradients_match_central_differences():
    rng = np.random.default_rng(9)
    params = [(w.astype(np.float64), b.astype(np.float64)) for w, b in R.random_params(rng, 3, 5, 3, 2)]
    x = rng.standard_normal((1, 4, 6, 3))
    t = (rng.random((1, 4, 6, 2)) < 0.5).astype(np.float64)
    m = (rng.random((1, 4, 6)) < 0.7).astype(np.float64)
    _, grads = R.loss_grad(params, x, t, m, quant=False)
    eps = 1e-6
    for L in range(len(params)):
        for which in (0, 1):
            arr = params[L][which]
            for idx in [tuple(rng.integers(0, s) for s in arr.shape) for _ in range(6)]:
                old = arr[idx]
                arr[idx] = old + eps
                lp, _ = R.loss_grad(params, x, t, m, quant=False)
                arr[idx] = old - eps
                lm, _ = R.loss_grad(params, x, t, m, quant=False)
                arr[idx] = old
                fd = (lp - lm) / (2 * eps)
                assert abs(fd - grads[L][which][idx]) < 1e-6 * max(1.0, abs(fd)), (L, which, idx)


def test_quantized_forward_stays_close_to_float():
    rng = np.random.default_rng(3)
    params = R.random_params(rng, 20, 32, 6, 4)
    x = rng.standard_normal((1, 8, 64, 20))
    zq, _ = R.forward(params, x, quant=True)
    zf, _ = R.forward(params, x, quant=False)
    assert np.abs(zq - zf).max() < 0.05 * (np.abs(zf).max() + 1)


def test_flatten_order_is_layerwise_w_then_b():
    p = [(np.arange(2 * 9 * 3).reshape(2, 3, 3, 3), np.array([100, 101])),
         (np.arange(1 * 9 * 2).reshape(1, 3, 3, 2) + 200, np.array([300]))]
    f = R.flatten(p)
    assert f.shape == (54 + 2 + 18 + 1,)
    assert f[0] == 0 and f[53] == 53 and f[54] == 100 and f[55] == 101 and f[56] == 200 and f[-1] == 300
