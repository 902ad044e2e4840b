"""Pins the float64 oracle of the learned denoiser.

The denoiser has no reference counterpart, so the oracle is pinned by what the
mathematics fixes: direct convolution loops, adjoint (dot-product) identities
of every resampling step, central finite differences of the whole network's
loss, Adam's closed form, BF16 rounding against torch's conversion, and
closed-form feature values.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import net as N


def test_layer_table_matches_the_package():
    import lucida_amd as L
    assert [tuple(l) for l in N.LAYERS] == [tuple(l) for l in L.NET_LAYERS]
    assert N.LEVEL == L.NET_LEVEL
    # skip concatenations: decoder input = upsampled low + skip channels
    assert N.LAYERS[8][0] == N.LAYERS[7][1] + N.LAYERS[5][1]
    assert N.LAYERS[10][0] == N.LAYERS[9][1] + N.LAYERS[3][1]
    assert N.LAYERS[12][0] == N.LAYERS[11][1] + N.LAYERS[1][1]
    assert N.param_count() == sum(9 * a * b + b for a, b, _ in N.LAYERS)


def test_conv_matches_direct_loops():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 5, 6, 3))
    W = rng.normal(size=(9, 3, 4))
    b = rng.normal(size=4)
    y = N.conv3x3(x, W, b)
    ref = np.zeros((2, 5, 6, 4))
    for n in range(2):
        for i in range(5):
            for j in range(6):
                for co in range(4):
                    s = b[co]
                    for dy in range(3):
                        for dx in range(3):
                            yy, xx = i + dy - 1, j + dx - 1
                            if 0 <= yy < 5 and 0 <= xx < 6:
                                s += x[n, yy, xx] @ W[3 * dy + dx, :, co]
                    ref[n, i, j, co] = s
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_conv_matches_torch_cross_correlation():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(1, 8, 8, 5))
    W = rng.normal(size=(9, 5, 7))
    tw = torch.tensor(W.reshape(3, 3, 5, 7)).permute(3, 2, 0, 1)  # (co, ci, ky, kx)
    ty = torch.nn.functional.conv2d(torch.tensor(x).permute(0, 3, 1, 2), tw, padding=1)
    np.testing.assert_allclose(N.conv3x3(x, W), ty.permute(0, 2, 3, 1).numpy(), rtol=1e-10, atol=1e-10)


def test_conv_gradients_are_adjoint():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(2, 4, 6, 3))
    W = rng.normal(size=(9, 3, 5))
    dy = rng.normal(size=(2, 4, 6, 5))
    dx, dW, db = N.conv3x3_grads(x, W, dy)
    # <conv(x), dy> is linear in x and in W: its gradients are dx and dW
    assert np.isclose((N.conv3x3(x, W) * dy).sum(), (x * dx).sum())
    assert np.isclose((N.conv3x3(x, W) * dy).sum(), (W * dW).sum())
    np.testing.assert_allclose(db, dy.sum(axis=(0, 1, 2)))


def test_resampling_adjoints():
    rng = np.random.default_rng(6)
    x = rng.normal(size=(2, 4, 8, 3))
    y = rng.normal(size=(2, 2, 4, 3))
    assert np.isclose((N.pool(x) * y).sum(), (x * N.pool_bwd(y)).sum())
    low = rng.normal(size=(2, 2, 4, 3))
    skip = rng.normal(size=(2, 4, 8, 2))
    d = rng.normal(size=(2, 4, 8, 5))
    dlow, dskip = N.upcat_bwd(d, 3)
    assert np.isclose((N.upcat(low, skip) * d).sum(), (low * dlow).sum() + (skip * dskip).sum())
    # average pool of a constant is the constant
    np.testing.assert_allclose(N.pool(np.full((1, 4, 4, 2), 3.5)), 3.5)


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = N.init_params(11)
    # random biases, so no unit starts dead everywhere
    layers = N.unpack(p.copy())
    for W, b in layers:
        b[:] = rng.normal(size=b.shape) * 0.05
    p = N.pack(layers)
    x = rng.normal(size=(1, 8, 16, 16)) * 0.5
    tgt = rng.normal(size=(1, 8, 16, 3)) * 0.5
    loss, g, _, _ = N.loss_and_grad(p, x, tgt)
    assert np.isfinite(loss) and g.shape == p.shape
    offs = np.cumsum([0] + [9 * a * b + b for a, b, _ in N.LAYERS])
    eps = 1e-6
    checked = 0
    for li in range(len(N.LAYERS)):
        ci, co, _ = N.LAYERS[li]
        idx = [offs[li] + rng.integers(0, 9 * ci * (3 if li == 14 else co)) // (3 if li == 14 else co) * co
               + rng.integers(0, 3 if li == 14 else co) for _ in range(3)]
        idx.append(offs[li] + 9 * ci * co + rng.integers(0, 3 if li == 14 else co))  # a bias
        for k in idx:
            pp, pm = p.copy(), p.copy()
            pp[k] += eps
            pm[k] -= eps
            fd = (N.loss_and_grad(pp, x, tgt)[0] - N.loss_and_grad(pm, x, tgt)[0]) / (2 * eps)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9, (li, k, fd, g[k])
            checked += 1
    assert checked == 4 * len(N.LAYERS)


def test_padding_channels_receive_no_gradient():
    """Feature channel 15 is zero, so its weights get zero gradient; output
    channels 3..15 are outside the loss."""
    rng = np.random.default_rng(8)
    p = N.init_params(2)
    x = np.zeros((1, 8, 16, 16))
    x[..., :N.INPUTS_USED] = rng.normal(size=(1, 8, 16, N.INPUTS_USED))
    tgt = rng.normal(size=(1, 8, 16, 3))
    _, g, _, _ = N.loss_and_grad(p, x, tgt)
    (dW0, _), *_, (dW14, db14) = N.unpack(g)
    assert np.all(dW0[:, N.INPUTS_USED:, :] == 0)
    assert np.all(dW14[:, :, 3:] == 0) and np.all(db14[3:] == 0)


def test_adam_closed_forms():
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    p = np.ones(4)
    m = v = np.zeros(4)
    p1, m1, v1 = N.adam(p, g, m, v, 1, 0.01)
    # first step: m_hat = g, v_hat = g^2, so the step is lr * g / (|g| + eps)
    np.testing.assert_allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    # a constant gradient keeps m_hat = g and v_hat = g^2 at every step
    p2, _, _ = N.adam(p1, g, m1, v1, 2, 0.01)
    np.testing.assert_allclose(p2, p1 - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.normal(size=4000) * 10.0 ** rng.integers(-6, 6, size=4000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 65504.0, 3.0e38]])
    t = torch.tensor(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(N.to_bf16(x.astype(np.float32)), t)


def test_feature_closed_forms():
    a = np.zeros((3, 16))
    ns = np.array([4.0, 1.0, 16.0])
    al = np.array([[0.5, 0.25, 0.8], [0.2, 0.2, 0.2], [1e-4, 0.5, 0.5]])
    a[:, 3] = ns
    a[:, 4:7] = al * ns[:, None]
    mean_rad = al * (np.e - 1.0)  # demodulated radiance e - 1 -> log1p = 1
    mean_rad[2, 0] = 1e-3 * (np.e - 1.0)  # albedo floored at 1e-3
    a[:, 0:3] = mean_rad * ns[:, None]
    a[:, 12:15] = mean_rad ** 2 * ns[:, None]  # zero variance
    a[:, 7] = np.array([1.0, 3.0, 0.0]) * ns
    a[:, 8:11] = np.array([0, 1, 0]) * ns[:, None]
    f = N.features(a)
    np.testing.assert_allclose(f[:, 0:3], 1.0, rtol=1e-12)
    np.testing.assert_allclose(f[:, 3:6], np.maximum(al, 1e-3))
    np.testing.assert_allclose(f[:, 7], 1.0)
    np.testing.assert_allclose(f[:, 9], [0.5, 0.75, 0.0])
    np.testing.assert_allclose(f[:, 10], 0.0, atol=1e-12)
    np.testing.assert_allclose(f[:, 11], np.log2(ns) / 16)
    assert np.all(f[:, 12:] == 0)
    # the a-trous channels: a base film (samples = 1) of radiance albedo * (e^2 - 1) -> 2
    base = np.zeros((3, 16))
    base[:, 3] = 1.0
    base[:, 0:3] = np.maximum(al, 1e-3) * (np.e ** 2 - 1.0)
    fb = N.features(a, base)
    np.testing.assert_allclose(fb[:, 12:15], 2.0, rtol=1e-12)
    np.testing.assert_array_equal(fb[:, :12], f[:, :12])
    assert np.all(fb[:, 15] == 0)
    # target of the input against itself is feature 0..2; compose inverts it
    # once the a-trous channels are added back (here: zero output + base 2)
    np.testing.assert_allclose(N.target(a, a), f[:, 0:3], rtol=1e-12)
    comp = N.compose(f[:, 0:3], a, f)
    np.testing.assert_allclose(comp[:, 0:3], mean_rad, rtol=1e-12)
    assert np.all(comp[:, 3] == 1.0)
    comp2 = N.compose(np.zeros((3, 16)), a, fb)
    np.testing.assert_allclose(comp2[:, 0:3], base[:, 0:3], rtol=1e-12)
