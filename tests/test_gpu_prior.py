"""The prior network's BF16 matrix-core kernels against the CPU restatement (oracle/prior_ref.py).

Tolerances (floating point): the oracle rounds to BF16 where the device does and sums in float64, the device sums in
FP32, so they differ by FP32 summation-order error and the rare BF16 rounding flip it causes downstream. Logits:
|dz| <= 2e-2 (1 + |z|). Gradients: relative L2 error <= 2e-2 per layer and cosine >= 0.999.
"""
import numpy as np
import pytest

from oracle import prior_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prior():
    from pcbroute_amd import prior
    return prior


def test_mfma_lane_layout_exact_on_small_integers(prior):
    """The 16x16x32 fragment layout the conv kernels assume: small integers multiply exactly in BF16 with FP32
    sums, and one-hot operands pick out a single row or column, so any transposed, permuted or misplaced fragment
    value changes the result."""
    rng = np.random.default_rng(0)
    a = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    b = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    assert np.array_equal(prior.mfma_check(a, b), a @ b)
    for (i, k) in ((3, 11), (15, 31), (0, 24), (9, 7)):
        e = np.zeros((16, 32), np.float32)
        e[i, k] = 1.0  # picks row k of b into row i
        assert np.array_equal(prior.mfma_check(e, b), e @ b)
        f = np.zeros((32, 16), np.float32)
        f[k, i] = 1.0  # picks column k of a into column i
        assert np.array_equal(prior.mfma_check(a, f), a @ f)


def test_mfma_rounds_inputs_to_bf16(prior):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((16, 32)).astype(np.float32)
    b = rng.standard_normal((32, 16)).astype(np.float32)
    want = R.bf16_round(a).astype(np.float64) @ R.bf16_round(b).astype(np.float64)
    np.testing.assert_allclose(prior.mfma_check(a, b), want, rtol=0, atol=1e-5)


def test_params_round_trip(prior):
    net = prior.PriorNet(20, 32, 6, 4, seed=3)
    assert net.num_params == 20 * 9 * 32 + 32 + 4 * (32 * 9 * 32 + 32) + 32 * 9 * 4 + 4
    rng = np.random.default_rng(2)
    v = R.flatten(R.random_params(rng, 20, 32, 6, 4))
    net.set_params(v)
    assert np.array_equal(net.get_params(), v)
    net.close()


def _net(prior, rng, cin, width, depth, cout):
    net = prior.PriorNet(cin, width, depth, cout, seed=1)
    params = R.random_params(rng, cin, width, depth, cout)
    net.set_params(R.flatten(params))
    return net, params


@pytest.mark.parametrize("shape", [(2, 7, 64, 20, 32, 6, 4), (1, 13, 128, 5, 16, 3, 2), (3, 1, 64, 1, 32, 1, 1),
                                   (1, 33, 192, 12, 24, 4, 4)])
def test_logits_match_oracle(prior, shape):
    n, h, w, cin, width, depth, cout = shape
    rng = np.random.default_rng(sum(shape))
    net, params = _net(prior, rng, cin, width, depth, cout)
    x = rng.standard_normal((n, h, w, cin)).astype(np.float32)
    got = net.logits(x)
    want, _ = R.forward(params, x, quant=True)
    assert np.all(np.abs(got - want) <= 2e-2 * (1 + np.abs(want))), np.abs(got - want).max()
    p = net.predict(x)
    np.testing.assert_allclose(p, 1 / (1 + np.exp(-got.astype(np.float64))), rtol=1e-5, atol=1e-6)
    net.close()


@pytest.mark.parametrize("shape", [(2, 9, 64, 20, 32, 6, 4), (1, 17, 128, 3, 8, 3, 2), (2, 4, 64, 2, 32, 2, 3)])
def test_loss_and_gradients_match_oracle(prior, shape):
    n, h, w, cin, width, depth, cout = shape
    rng = np.random.default_rng(7 + sum(shape))
    net, params = _net(prior, rng, cin, width, depth, cout)
    x = rng.standard_normal((n, h, w, cin)).astype(np.float32)
    t = (rng.random((n, h, w, cout)) < 0.3).astype(np.float32)
    m = (rng.random((n, h, w)) < 0.8).astype(np.float32)
    loss, g = net.loss_grad(x, t, m)
    lref, gref = R.loss_grad(params, x, t, m, quant=True)
    assert abs(loss - lref) <= 1e-3 * (1 + abs(lref))
    got = net.unpack(g)
    for L, ((gw, gb), (rw, rb)) in enumerate(zip(got, gref)):
        for a, b in ((gw, rw), (gb, rb)):
            a = a.astype(np.float64).ravel()
            b = b.ravel()
            nb = np.linalg.norm(b)
            assert np.linalg.norm(a - b) <= 2e-2 * nb + 1e-7, (L, np.linalg.norm(a - b) / max(nb, 1e-30))
            if nb > 1e-6:
                assert a @ b / (np.linalg.norm(a) * nb) >= 0.999, L
    net.close()


def test_mask_excludes_pixels(prior):
    rng = np.random.default_rng(11)
    net, params = _net(prior, rng, 4, 16, 2, 2)
    x = rng.standard_normal((1, 6, 64, 4)).astype(np.float32)
    t = (rng.random((1, 6, 64, 2)) < 0.5).astype(np.float32)
    m = np.zeros((1, 6, 64), np.float32)
    loss, g = net.loss_grad(x, t, m)
    assert loss == 0.0 and not np.any(g)
    m[0, 2, 5] = 1
    loss, g = net.loss_grad(x, t, m)
    lref, _ = R.loss_grad(params, x, t, m, quant=True)
    assert abs(loss - lref) < 1e-3 * (1 + lref) and np.any(g)
    net.close()


def test_training_lowers_the_loss(prior):
    # a learnable map: the target is a 3x3 dilation of input channel 0 being positive
    rng = np.random.default_rng(5)
    net = prior.PriorNet(3, 16, 3, 1, seed=4)
    x = rng.standard_normal((4, 16, 64, 3)).astype(np.float32)
    pos = x[..., 0] > 1.0
    t = np.zeros((4, 16, 64, 1), np.float32)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            t[..., 0] = np.maximum(t[..., 0], np.roll(np.roll(pos, dy, 1), dx, 2))
    m = np.ones((4, 16, 64), np.float32)
    first = net.train_step(x, t, m, lr=3e-3)
    for _ in range(150):
        last = net.train_step(x, t, m, lr=3e-3)
    assert last < 0.5 * first, (first, last)
    net.close()


def test_bad_shapes_raise(prior):
    net = prior.PriorNet(4, 8, 2, 1)
    with pytest.raises(ValueError):
        net.logits(np.zeros((1, 4, 60, 4), np.float32))  # width not a multiple of 64
    lib = prior.library()
    x = np.zeros(4 * 60 * 4, np.float32)
    out = np.zeros(4 * 60, np.float32)
    assert lib.pcbprior_logits(net.h, prior._p(x), 1, 4, 60, prior._p(out)) == -1
    assert b"64" in lib.pcbprior_last_error()
    with pytest.raises(prior.BackendError):
        prior.PriorNet(40, 8, 2, 1)  # more channels than a layer carries
    net.close()


def test_resident_training_matches_host_batches(prior):
    """A step on the resident set equals the step on the same batch passed from the host; forward_resident equals
    logits; the accounting counts every conv launch with the algorithmic FLOPs."""
    rng = np.random.default_rng(21)
    cin, width, depth, cout = 6, 16, 3, 2
    params = R.flatten(R.random_params(rng, cin, width, depth, cout))
    x = rng.standard_normal((5, 8, 64, cin)).astype(np.float32)
    t = (rng.random((5, 8, 64, cout)) < 0.3).astype(np.float32)
    m = (rng.random((5, 8, 64)) < 0.9).astype(np.float32)
    a = prior.PriorNet(cin, width, depth, cout)
    b = prior.PriorNet(cin, width, depth, cout)
    a.set_params(params)
    b.set_params(params)
    idx = np.array([3, 0, 4], np.int32)
    b.dataset(x, t, m)
    np.testing.assert_allclose(b.forward_resident(idx), a.logits(x[idx]), rtol=1e-6, atol=1e-6)
    b.stats(enable=True, reset=True)
    la = a.train_step(x[idx], t[idx], m[idx], lr=1e-2)
    lb = b.train_resident(idx, lr=1e-2)
    st = b.stats(enable=False, reset=True)
    assert abs(la - lb) <= 1e-6 * (1 + abs(la))
    # the weight gradients sum FP32 partials with atomics, so the two steps differ in summation order only; Adam turns
    # a near-zero gradient into a full step of either sign, so a few parameters may differ by up to 2 lr
    pa, pb = a.get_params(), b.get_params()
    off = np.abs(pa - pb) > 1e-5 * (1 + np.abs(pa))
    assert off.mean() < 0.01 and np.abs(pa - pb).max() <= 2.01e-2
    assert st["conv_launches"] == depth + depth - 1 and st["wgrad_launches"] == depth
    npx = 3 * 8 * 64
    fwd = 2 * npx * 9 * (cin * width + width * width + width * cout)
    assert abs(st["conv_flop"] - fwd - (fwd - 2 * npx * 9 * cin * width)) < 1 and abs(st["wgrad_flop"] - fwd) < 1
    assert st["conv_ms"] > 0 and st["wgrad_ms"] > 0
    with pytest.raises(prior.BackendError):
        b.train_resident(np.array([5], np.int32))  # out of range
    a.close()
    b.close()


def test_padded_batch_equals_each_image_alone(prior):
    """pcbprior_predict_valid: images of different sizes padded into one batch give exactly the probabilities each
    gives alone (the valid extent zeroes every layer's activations outside it, as the zero padding of a lone image)."""
    rng = np.random.default_rng(31)
    net, _ = _net(prior, rng, 22, 32, 6, 4)
    xs = [rng.random((h, w, 22)).astype(np.float32) for h, w in ((13, 64), (40, 128), (7, 192), (40, 64))]
    alone = [net.predict(x[None])[0] for x in xs]
    many = net.predict_many(xs)
    for a, b in zip(alone, many):
        assert a.shape == b.shape and np.array_equal(a, b)
    lib = prior.library()
    out = np.zeros(4 * 64 * 4, np.float32)
    bad = np.array([5, 60], np.int32)  # columns not a multiple of 64
    assert lib.pcbprior_predict_valid(net.h, prior._p(np.zeros(4 * 64 * 22, np.float32)), 1, 4, 64,
                                      bad.ctypes.data_as(prior.C.POINTER(prior.C.c_int32)), prior._p(out)) == -1
    net.close()
