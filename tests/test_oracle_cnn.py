"""Pins of the FP64 U-Net oracle (CPU only): the convolution against closed
forms (the identity tap, a shifting tap, a constant kernel on a ramp), the
whole network's analytic gradient against central finite differences of its
own loss, and the parameter count against the architecture's formula."""
import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(77)


def test_conv_identity_and_shift_taps():
    B, H, W, C = 2, 5, 7, 4
    X = RNG.standard_normal((B, H, W, C))
    Wt = np.zeros((C, 9, C))
    for c in range(C):
        Wt[c, 4, c] = 1.0                     # centre tap t = 4 (dy = dx = 0)
    np.testing.assert_array_equal(oracle.conv3x3(X, Wt), X)
    Wt = np.zeros((C, 9, C))
    for c in range(C):
        Wt[c, 5, c] = 1.0                     # t = 5: dy = 0, dx = +1 -> Y[x] = X[x + 1]
    Y = oracle.conv3x3(X, Wt, b=np.arange(C, dtype=np.float64))
    np.testing.assert_allclose(Y[:, :, :-1, :] - np.arange(C), X[:, :, 1:, :], atol=1e-14)
    np.testing.assert_array_equal(Y[:, :, -1, :], np.broadcast_to(np.arange(C, dtype=np.float64), (B, H, C)))
    Wt = np.zeros((C, 9, C))
    Wt[:, 1, 0] = 2.0                         # t = 1: dy = -1, dx = 0, from input channel 0 into every output
    Y = oracle.conv3x3(X, Wt)
    np.testing.assert_array_equal(Y[:, 1:, :, 2], 2.0 * X[:, :-1, :, 0])
    assert np.all(Y[:, 0] == 0.0)


def test_conv_constant_kernel_on_a_ramp():
    """All-ones 3x3 kernel on a linear ramp: 9 x the ramp inside, the partial
    sums of the zero-padded border outside."""
    H, W = 6, 8
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    X = (0.5 + 0.25 * xs - 0.125 * ys)[None, :, :, None]
    Y = oracle.conv3x3(X, np.ones((1, 9, 1)))[0, :, :, 0]
    np.testing.assert_allclose(Y[1:-1, 1:-1], 9.0 * X[0, 1:-1, 1:-1, 0], atol=1e-13)
    want00 = X[0, 0, 0, 0] + X[0, 0, 1, 0] + X[0, 1, 0, 0] + X[0, 1, 1, 0]
    assert abs(Y[0, 0] - want00) < 1e-13


@pytest.mark.parametrize("levels,cpad", [(1, 32), (2, 32), (3, 64)])
def test_parameter_count(levels, cpad):
    ch = [min(32 << l, 256) for l in range(levels)]
    n = 0
    cin = cpad
    for l in range(levels):
        n += ch[l] * 9 * cin + ch[l] + ch[l] * 9 * ch[l] + ch[l]
        cin = ch[l]
    for l in range(levels - 2, -1, -1):
        n += ch[l] * 9 * ch[l + 1] + ch[l] + ch[l] * 9 * ch[l] + ch[l]
    n += 32 * 9 * ch[0] + 32
    assert oracle.cnn_num_params(levels, 32, cpad) == n


def test_network_gradient_matches_finite_differences():
    levels, H, W, B, cpad = 2, 4, 6, 1, 32
    n = oracle.cnn_num_params(levels, 32, cpad)
    params = 0.2 * RNG.standard_normal(n)
    x = RNG.standard_normal((B, H, W, cpad))
    base = RNG.standard_normal((B, 3, H, W))
    target = RNG.standard_normal((B, 3, H, W))
    loss, pred, g = oracle.cnn_forward_backward(levels, 32, x, params, base, target)
    assert np.isfinite(loss) and loss > 0
    # the head's unused output channels carry no gradient
    head_w = 32 * 9 * 32
    head = params.size - head_w - 32
    gh = g[head:head + head_w].reshape(32, 9, 32)
    assert np.all(gh[3:] == 0.0) and np.any(gh[:3] != 0.0)
    # central differences on parameters of every layer (largest-gradient ones
    # and random ones): the loss is piecewise quadratic, h small
    idx = np.concatenate([np.argsort(-np.abs(g))[:20], RNG.choice(n, 40, replace=False)])
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        l1, _, _ = oracle.cnn_forward_backward(levels, 32, x, p1, base, target, grads=False)
        l2, _, _ = oracle.cnn_forward_backward(levels, 32, x, p2, base, target, grads=False)
        fd = (l1 - l2) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])


def test_prediction_is_base_plus_head():
    levels, H, W, B, cpad = 2, 4, 4, 2, 32
    n = oracle.cnn_num_params(levels, 32, cpad)
    params = np.zeros(n)
    params[-32:-29] = [0.1, -0.2, 0.3]   # head bias of the 3 used channels
    x = RNG.standard_normal((B, H, W, cpad))
    base = RNG.standard_normal((B, 3, H, W))
    loss, pred, _ = oracle.cnn_forward_backward(levels, 32, x, params, base, base, grads=False)
    np.testing.assert_allclose(pred - base, np.array([0.1, -0.2, 0.3])[None, :, None, None] * np.ones_like(base),
                               atol=1e-15)
    assert abs(loss - (0.01 + 0.04 + 0.09) / 3) < 1e-14
