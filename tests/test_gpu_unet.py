"""Learned initial guess on the GPU: the MFMA operand layout, the network's forward pass,
loss and every parameter gradient against the CPU oracle (oracle/unet_oracle.py with the
GPU's bf16 rounding points), and training on solver-generated pairs with the solver run
from the network's x0."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import unet_oracle as UO  # noqa: E402

pytestmark = pytest.mark.gpu


def test_mfma_operand_layout(gpu):
    X = gpu
    rng = np.random.default_rng(0)
    a = rng.integers(-8, 9, size=(16, 32)).astype(np.float32)
    b = rng.integers(-8, 9, size=(32, 16)).astype(np.float32)
    assert np.array_equal(X.mfma_selftest(a, b), a @ b)  # small integers: exact


def layer_slices(channels):
    shapes = []
    for l in range(len(channels)):
        shapes += [(1 if l == 0 else channels[l - 1], channels[l]), (channels[l], channels[l])]
    for l in range(len(channels) - 1):
        shapes += [(channels[l + 1], channels[l]), (channels[l], channels[l])]
    out, o = [], 0
    for a, b in shapes:
        out.append(slice(o, o + 27 * a * b))
        o += 27 * a * b
        out.append(slice(o, o + b))
        o += b
    out.append(slice(o, o + channels[0] + 1))
    return out


@pytest.mark.parametrize("channels,size,batch", [((16, 32, 48), 8, 2),
                                                 ((16, 32, 48, 64), 8, 1),
                                                 ((16,), 6, 3),
                                                 # 64^3: the 16-channel convolutions run
                                                 # on LDS-staged tiles
                                                 ((16, 32), 64, 1)])
def test_loss_and_gradients_match_oracle(gpu, channels, size, batch):
    X = gpu
    net = X.UNet(channels, seed=7)
    p = net.get_params()
    # non-zero biases so that every bias path is exercised
    rng = np.random.default_rng(1)
    for sl in layer_slices(channels)[1::2]:
        p[sl] = rng.standard_normal(sl.stop - sl.start).astype(np.float32) * 0.1
    net.set_params(p)
    x = rng.standard_normal((batch, size, size, size)).astype(np.float32)
    t = rng.standard_normal((batch, size, size, size)).astype(np.float32)
    loss, g, y = net.loss_grad(x, t)
    ref = UO.UNet(channels, p.astype(np.float64))
    rloss, rg, ry = ref.loss_grad(x.astype(np.float64), t.astype(np.float64), use_bf16=True)
    assert abs(loss - rloss) <= 1e-3 * rloss
    assert np.linalg.norm(y - ry) <= 1e-2 * np.linalg.norm(ry)
    for sl in layer_slices(channels):
        gg, rr = g[sl], rg[sl]
        assert np.linalg.norm(gg - rr) <= 3e-2 * np.linalg.norm(rr) + 1e-9, sl


def test_training_and_solver_from_learned_x0(gpu):
    X = gpu
    N, nv = 32, 8
    row, col, val, _ = X.varcoef7(N, sigma=1.0, seed=3)
    levels = X.Amg(row, col, val, agg_num_levels=2, num_paths=2, coarse_size=500).export()
    prm = X.default_params(method="MergedPBiCGStab", smoother="Chebyshev", poly_order=4,
                           rel_tol=1e-8, max_iters=100)
    S = X.Solver(X.Hier(levels), prm, nv)
    pairs = []
    for q in range(6):
        b = X.rhs_bumps(100 + q, N, N, N, nv)
        db, dx = X.DeviceArray.from_host(b), X.DeviceArray(N ** 3 * nv)
        S.solve_device(db, dx)
        assert S.stats()[1].all()
        pairs.append((db, dx))
    net = X.UNet((16, 32, 48, 64), seed=1)
    hist = net.train(N, N, N, pairs, nv, crop=16, batch=8, steps=200, lr=2e-3, seed=5)
    assert np.isfinite(hist).all()
    assert hist[-20:].mean() < 0.5 * hist[:20].mean()
    # held-out right-hand sides: the error of x0 against the solver's solution (the
    # residual b - A x0 is not the measure: A scales the high-frequency part of the error
    # by up to 12/h^2, and that part is what the smoother removes first)
    b = X.rhs_bumps(999, N, N, N, nv)
    db = X.DeviceArray.from_host(b)
    dx0 = X.DeviceArray(N ** 3 * nv)
    net.infer(N, N, N, nv, db, dx0)
    x0 = dx0.to_host().reshape(-1, nv)
    it = {}
    sol = {}
    for name, start in (("zero", None), ("learned", x0.reshape(-1))):
        dx = X.DeviceArray.from_host(start) if start is not None else X.DeviceArray(N ** 3 * nv)
        S.solve_device(db, dx)
        st = S.stats()
        assert st[1].all()
        it[name] = st[0]
        sol[name] = dx.to_host().reshape(-1, nv)
    xs = sol["zero"]
    err = np.linalg.norm(x0 - xs, axis=0) / np.linalg.norm(xs, axis=0)
    # a short training run (200 steps on 16^3 crops of 48 pairs) already removes an
    # iteration; the error of x0 is reported, not bounded
    print("relative error of x0 per RHS:", np.round(err, 3), "iterations:", it)
    assert np.isfinite(err).all()
    assert it["learned"] <= it["zero"], it
    # both starts reach the same solution to the solver's tolerance
    assert (np.abs(sol["learned"] - xs).max(0) / np.abs(xs).max(0)).max() < 1e-6


def test_fused_upsampling_inference_is_identical(gpu):
    # inference reads the decoder's upsampled inputs through the convolution (no
    # upsampled tensor is written); the arithmetic is that of the materialised path
    X = gpu
    net = X.UNet((16, 32, 48), seed=3)
    net.output_scale = 0.5
    for n, nv in ((16, 3), (64, 2)):
        b = X.DeviceArray.from_host(np.random.default_rng(n).standard_normal(n ** 3 * nv))
        outs = []
        for fuse in (0, 1):
            X.set_option("unet_fuse_up", fuse)
            x = X.DeviceArray(n ** 3 * nv)
            net.infer(n, n, n, nv, b, x)
            outs.append(x.to_host())
        X.set_option("unet_fuse_up", 1)
        assert np.isfinite(outs[0]).all() and np.abs(outs[0]).max() > 0
        assert np.array_equal(outs[0], outs[1])


def test_tiled_convolution_is_identical(gpu):
    # forward convolutions with 16 or 32 output channels on volumes with w % 64 == h % 4 ==
    # d % 2 == 0 (48 channels: w % 32 == 0) read their input from LDS-staged tiles (plain,
    # and through the fused upsampling); products, k order and epilogue are those of the
    # per-tap gathers, so the result is bit-identical
    X = gpu
    try:
        # level-1 tiles (32 output channels) on the 128 x 32 x 16 grid: 16 -> 32, 32 -> 32
        # (2-line tiles) and the upsampled 48 -> 32 and 32 -> 32; level-2 tiles (48 output
        # channels, lines of 32 voxels): 32 -> 48, 48 -> 48 and the upsampled 64 -> 48
        for ch, fuse in (((16, 32, 48), 1), ((16, 32, 48), 0), ((16, 16, 32), 1),
                         ((16, 32, 32), 1), ((16, 32, 48, 64), 1)):
            net = X.UNet(ch, seed=5)
            net.output_scale = 0.5
            X.set_option("unet_fuse_up", fuse)
            for (nx, ny, nz), nv in (((64, 64, 64), 2), ((128, 32, 16), 3)):
                V = nx * ny * nz
                b = X.DeviceArray.from_host(np.random.default_rng(nx + nv).standard_normal(V * nv))
                outs = []
                for tile in (0, 1):
                    X.set_option("unet_conv_tile", tile)
                    x = X.DeviceArray(V * nv)
                    net.infer(nx, ny, nz, nv, b, x)
                    outs.append(x.to_host())
                assert np.isfinite(outs[0]).all() and np.abs(outs[0]).max() > 0
                assert np.array_equal(outs[0], outs[1]), (ch, fuse, nx, ny, nz)
    finally:
        X.set_option("unet_conv_tile", 1)
        X.set_option("unet_fuse_up", 1)


@pytest.mark.parametrize("nv", [2, 3])
def test_inference_matches_oracle_forward(gpu, nv):
    # x0[:, k] = rms(b[:, k]) * o * net(b[:, k] / rms(b[:, k])) on a grid the LDS-staged
    # convolution tiles cover (the single-channel input layer included); nv = 2 takes the
    # vector head, nv = 3 the per-element one
    X = gpu
    channels = (16, 32)
    net = X.UNet(channels, seed=9)
    p = net.get_params()
    rng = np.random.default_rng(4)
    for sl in layer_slices(channels)[1::2]:
        p[sl] = rng.standard_normal(sl.stop - sl.start).astype(np.float32) * 0.1
    net.set_params(p)
    net.output_scale = 0.7
    nx, ny, nz = 64, 16, 8
    V = nx * ny * nz
    b = rng.standard_normal((V, nv))
    x0 = X.DeviceArray(V * nv)
    net.infer(nx, ny, nz, nv, X.DeviceArray.from_host(b.ravel()), x0)
    got = x0.to_host().reshape(V, nv)
    scale = np.sqrt((b ** 2).mean(0))
    inp = (b / scale).T.reshape(nv, nz, ny, nx)
    ref = UO.UNet(channels, p.astype(np.float64))
    _, _, ry = ref.loss_grad(inp, np.zeros_like(inp), use_bf16=True)
    exp = (scale[:, None] * 0.7 * ry.reshape(nv, V)).T
    assert np.linalg.norm(got - exp) <= 1e-2 * np.linalg.norm(exp)
