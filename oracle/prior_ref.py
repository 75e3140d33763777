"""CPU restatement of the learned corridor prior's network (numpy), the checker of the BF16 matrix-core kernels.

TEST INFRASTRUCTURE ONLY: imported by tests/ and __graft_entry__.smoke() as the checker of the HIP kernels; the
product package never imports it.

The network is the one pcbroute_amd.prior.PriorNet runs: `depth` conv3x3 layers ("same" zero padding, stride 1),
ReLU between layers, per-pixel logits out, trained with the mean binary cross-entropy over masked pixels. It takes
the place of pcbroute's RouteNet call (src/mlinfer.cpp:834-856), whose ONNX graph is third-party arithmetic that
nothing here can run, so the network itself is this project's own and the textbook definitions pin it:
- conv3x3 as cross-correlation, checked against torch.nn.functional.conv2d and its autograd on the CPU;
- gradients checked by central finite differences in float64;
- BF16 rounding (round to nearest, ties to even) checked bit by bit against torch's float32 -> bfloat16 cast.

`quant=True` reproduces where the device rounds to BF16: the network input, the weights, each hidden activation (after
bias and ReLU), the loss gradient at the logits, and each back-propagated activation gradient. Products and sums
stay in float64 here, FP32 on the device, so the two agree to FP32 summation-order error plus the rare BF16 rounding
flip that error causes. `quant=False` is plain float64 arithmetic (the textbook network).

Layouts: x (n, h, w, ci) pixel-major; a layer's weights w (co, 3, 3, ci) = w[co][ky][kx][ci], biases b (co,).
"""
import numpy as np


def bf16_round(a):
    """float32 -> nearest BF16 (ties to even), returned as float32; NaN stays NaN."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    nan &= (u & 0x7FFFFF) != 0
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(a):
    return (bf16_round(a).view(np.uint32) >> 16).astype(np.uint16)


def _q(a, quant):
    return bf16_round(a).astype(np.float64) if quant else np.asarray(a, np.float64)


def im2col(x):
    """(n, h, w, c) -> (n, h, w, 9 c): the 3x3 neighbourhood of every pixel, zero outside, order (ky, kx, c)."""
    n, h, w, c = x.shape
    xp = np.zeros((n, h + 2, w + 2, c), x.dtype)
    xp[:, 1:-1, 1:-1] = x
    return np.concatenate([xp[:, ky:ky + h, kx:kx + w] for ky in range(3) for kx in range(3)], axis=3)


def conv3x3(x, w, b=None):
    """y[n, y, x, co] = sum over ky, kx, ci of w[co, ky, kx, ci] * x[n, y + ky - 1, x + kx - 1, ci] (+ b[co])."""
    co = w.shape[0]
    y = im2col(x) @ w.reshape(co, -1).T
    return y + b if b is not None else y


def forward(params, x, quant=True):
    """Logits (n, h, w, out_ch) and the activations entering each layer (acts[0] = the network input)."""
    a = _q(x, quant)
    acts = []
    for L, (w, b) in enumerate(params):
        acts.append(a)
        z = conv3x3(a, _q(w, quant), np.asarray(b, np.float64))
        if L < len(params) - 1:
            a = _q(np.maximum(z, 0.0), quant)
        else:
            return z, acts


def bce_with_logits(z, t, m):
    """Mean over the masked pixels and all channels of -t log s(z) - (1 - t) log(1 - s(z)), and its gradient."""
    m = np.asarray(m, np.float64)[..., None] != 0
    cnt = m.sum() * z.shape[-1]
    inv = 1.0 / cnt if cnt else 0.0
    l = np.maximum(z, 0) - z * t + np.log1p(np.exp(-np.abs(z)))
    g = (1.0 / (1.0 + np.exp(-z)) - t) * inv
    return float((l * m).sum() * inv), np.where(m, g, 0.0)


def loss_grad(params, x, t, m, quant=True):
    """(loss, [(dw, db)] per layer) of the mean BCE, back-propagated through the network."""
    z, acts = forward(params, x, quant)
    loss, dz = bce_with_logits(z, np.asarray(t, np.float64), m)
    dy = _q(dz, quant)
    grads = [None] * len(params)
    for L in range(len(params) - 1, -1, -1):
        a = acts[L]
        w = _q(params[L][0], quant)
        co = w.shape[0]
        cols = im2col(a).reshape(-1, a.shape[-1] * 9)
        grads[L] = ((dy.reshape(-1, co).T @ cols).reshape(w.shape), dy.reshape(-1, co).sum(0))
        if L > 0:
            # d/da of the correlation: the correlation of dy with the kernel rotated 180 degrees, channels swapped
            wr = w[:, ::-1, ::-1, :].transpose(3, 1, 2, 0)
            dx = conv3x3(dy, wr)
            dy = np.where(a > 0, _q(dx, quant), 0.0)
    return loss, grads


def flatten(pairs):
    """[(w, b)] -> the flat order of pcbprior_get_params: per layer w[co][ky][kx][ci] then b[co]."""
    return np.concatenate([np.concatenate([np.ravel(w), np.ravel(b)]) for w, b in pairs]).astype(np.float32)


def random_params(rng, in_ch, width, depth, out_ch, scale=1.0):
    out = []
    for L in range(depth):
        ci = in_ch if L == 0 else width
        co = out_ch if L == depth - 1 else width
        lim = scale * np.sqrt(6.0 / (9 * ci))
        out.append((rng.uniform(-lim, lim, (co, 3, 3, ci)).astype(np.float32),
                    rng.uniform(-0.1, 0.1, co).astype(np.float32)))
    return out
