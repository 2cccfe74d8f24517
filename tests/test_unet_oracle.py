"""The network oracle (oracle/unet_oracle.py) pinned by mathematics: every backward
operator is the adjoint of its forward operator, and the whole-network gradient matches
central finite differences of the loss (float64, no bf16 rounding)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import unet_oracle as U  # noqa: E402


def test_conv_backward_operators_are_adjoints():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 6, 4, 3))
    w = rng.standard_normal((27, 3, 5))
    g = rng.standard_normal((2, 4, 6, 4, 5))
    lhs = np.sum(U.conv(x, w, None) * g)
    assert np.isclose(lhs, np.sum(x * U.conv_data(g, w)), rtol=1e-12)
    gw, gb = U.conv_wgrad(x, g)
    assert np.isclose(lhs, np.sum(w * gw), rtol=1e-12)
    assert np.allclose(gb, g.reshape(-1, 5).sum(0))


def test_conv_against_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 3, 4, 5, 2))
    w = rng.standard_normal((27, 2, 3))
    out = U.conv(x, w, np.zeros(3))
    z, y, xx = 1, 2, 0  # a voxel on the x boundary
    ref = np.zeros(3)
    for t, (dz, dy, dx) in enumerate(U.TAPS):
        if 0 <= z + dz < 3 and 0 <= y + dy < 4 and 0 <= xx + dx < 5:
            ref += x[0, z + dz, y + dy, xx + dx] @ w[t]
    assert np.allclose(out[0, z, y, xx], ref)


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 4, 4, 6, 3))
    g = rng.standard_normal((2, 2, 2, 3, 3))
    assert np.isclose(np.sum(U.pool(x) * g), np.sum(x * U.pool_bwd(g)))
    y = rng.standard_normal((2, 2, 2, 3, 3))
    h = rng.standard_normal((2, 4, 4, 6, 3))
    assert np.isclose(np.sum(U.up(y) * h), np.sum(y * U.up_bwd(h)))


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-40])
    r = U.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] - (-3.140625)) < 1e-12
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8 + 1e-38)


@pytest.mark.parametrize("channels", [(16, 16), (16, 32, 16)])
def test_network_gradient_matches_finite_differences(channels):
    rng = np.random.default_rng(3)
    # parameter count from the layer shapes
    shapes = []
    for l in range(len(channels)):
        shapes += [(1 if l == 0 else channels[l - 1], channels[l]), (channels[l], channels[l])]
    for l in range(len(channels) - 1):
        shapes += [(channels[l + 1], channels[l]), (channels[l], channels[l])]
    n = sum(27 * a * b + b for a, b in shapes) + channels[0] + 1
    p = rng.standard_normal(n) * 0.2
    net = U.UNet(channels, p)
    s = 2 ** (len(channels) - 1) * 2
    x = rng.standard_normal((2, s, s, s))
    t = rng.standard_normal((2, s, s, s))
    loss, g, _ = net.loss_grad(x, t)
    idx = rng.choice(n, size=25, replace=False)
    eps = 1e-6
    for i in idx:
        pp = p.copy()
        pp[i] += eps
        lp = U.UNet(channels, pp).loss_grad(x, t)[0]
        pp[i] -= 2 * eps
        lm = U.UNet(channels, pp).loss_grad(x, t)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])
