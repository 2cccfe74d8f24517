"""CPU oracle of the learned-initial-guess network (test infrastructure only: used by
tests/ to check the GPU kernels, never by the library).

numpy restatement of the U-Net of include/xamg_unet.h: forward pass, loss
0.5 * mean((y - t)^2) and the gradient of every parameter, in the parameter order of
xh_unet_get_params.  With bf16=True the values the GPU stores in bf16 (weights in their
GEMM layouts, every activation and activation gradient) are rounded to bf16 at the same
points, accumulation stays in float64; with bf16=False everything is float64 (the form the
finite-difference checks pin).
"""
import numpy as np


def bf16(a):
    """round float32 values to the nearest bfloat16 (ties to even), returned as float64"""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shift(x, dz, dy, dx):
    """y[n, z, y, x] = x[n, z+dz, y+dy, x+dx] with zeros outside ([N, D, H, W, C])"""
    out = np.zeros_like(x)
    N, D, H, W, _ = x.shape
    zs, ys, xs = slice(max(0, -dz), min(D, D - dz)), slice(max(0, -dy), min(H, H - dy)), \
        slice(max(0, -dx), min(W, W - dx))
    zt, yt, xt = slice(max(0, dz), min(D, D + dz)), slice(max(0, dy), min(H, H + dy)), \
        slice(max(0, dx), min(W, W + dx))
    out[:, zs, ys, xs] = x[:, zt, yt, xt]
    return out


def conv(x, w, b):
    """3x3x3 zero-padded convolution: out[m, co] = sum_t x[m + off_t] @ w[t] + b"""
    out = np.zeros(x.shape[:4] + (w.shape[2],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += shift(x, dz, dy, dx) @ w[t]
    return out + (0.0 if b is None else b)


def conv_data(g, w):
    """gradient w.r.t. the input of conv: gin[m] = sum_t g[m - off_t] @ w[t].T"""
    out = np.zeros(g.shape[:4] + (w.shape[1],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += shift(g, -dz, -dy, -dx) @ w[t].T
    return out


def conv_wgrad(x, g):
    """gw[t] = sum_m x[m + off_t]^T g[m], gb = sum_m g[m]"""
    ci, co = x.shape[-1], g.shape[-1]
    gw = np.zeros((27, ci, co))
    g2 = g.reshape(-1, co)
    for t, (dz, dy, dx) in enumerate(TAPS):
        gw[t] = shift(x, dz, dy, dx).reshape(-1, ci).T @ g2
    return gw, g2.sum(0)


def pool(x):
    N, D, H, W, C = x.shape
    return x.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_bwd(g):
    N, D, H, W, C = g.shape
    return g.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


class UNet:
    """parameters from a flat vector in the xh_unet_get_params order"""

    def __init__(self, channels, flat):
        self.ch = list(channels)
        L = self.L = len(channels)
        self.shapes = []
        for l in range(L):
            self.shapes.append((1 if l == 0 else channels[l - 1], channels[l]))
            self.shapes.append((channels[l], channels[l]))
        for l in range(L - 1):
            self.shapes.append((channels[l + 1], channels[l]))
            self.shapes.append((channels[l], channels[l]))
        self.set_flat(np.asarray(flat, np.float64))

    def nparams(self):
        return sum(27 * a * b + b for a, b in self.shapes) + self.ch[0] + 1

    def set_flat(self, p):
        self.W, self.B = [], []
        o = 0
        for a, b in self.shapes:
            self.W.append(p[o:o + 27 * a * b].reshape(27, a, b))
            o += 27 * a * b
            self.B.append(p[o:o + b])
            o += b
        self.hw = p[o:o + self.ch[0]]
        self.hb = p[o + self.ch[0]]
        assert o + self.ch[0] + 1 == len(p)

    def loss_grad(self, x_in, tgt, use_bf16=False):
        """x_in, tgt: [N, D, H, W]; returns (loss, flat gradient, y)"""
        q = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
        L = self.L
        Wq = [q(w) for w in self.W]
        relu = lambda a: np.maximum(a, 0.0)  # noqa: E731
        enc, dec = list(range(2 * L)), list(range(2 * L, 2 * L + 2 * (L - 1)))
        ins, As, Ss, Us, Ds, Hs = [q(x_in[..., None])], [], [], {}, {}, {}
        for l in range(L):
            a = q(relu(conv(ins[l], Wq[enc[2 * l]], self.B[enc[2 * l]])))
            s = q(relu(conv(a, Wq[enc[2 * l + 1]], self.B[enc[2 * l + 1]])))
            As.append(a)
            Ss.append(s)
            if l + 1 < L:
                ins.append(q(pool(s)))
        h = Ss[L - 1]
        for l in range(L - 2, -1, -1):
            u = q(up(h))
            d = q(relu(conv(u, Wq[dec[2 * l]], self.B[dec[2 * l]]) + Ss[l]))
            h = q(relu(conv(d, Wq[dec[2 * l + 1]], self.B[dec[2 * l + 1]])))
            Us[l], Ds[l], Hs[l] = u, d, h
        h0 = h
        y = h0 @ self.hw + self.hb
        M = y.size
        diff = y - tgt
        loss = 0.5 * np.sum(diff * diff) / M
        gy = diff / M
        gW = [None] * len(self.W)
        gB = [None] * len(self.W)
        ghw = np.tensordot(gy, h0, axes=4)  # sum over voxels
        ghb = gy.sum()
        gh = q(gy[..., None] * self.hw)
        gs = {}
        for l in range(L - 1):
            g = gh * (Hs[l] > 0)
            gW[dec[2 * l + 1]], gB[dec[2 * l + 1]] = conv_wgrad(Ds[l], g)
            gd = q(conv_data(g, Wq[dec[2 * l + 1]])) * (Ds[l] > 0)
            gW[dec[2 * l]], gB[dec[2 * l]] = conv_wgrad(Us[l], gd)
            gs[l] = gd
            gu = q(conv_data(gd, Wq[dec[2 * l]]))
            gh = q(up_bwd(gu))
        gs[L - 1] = gh
        gin = None
        for l in range(L - 1, -1, -1):
            g = gs[l]
            if l + 1 < L:
                g = q(g + q(pool_bwd(gin)))
            g = g * (Ss[l] > 0)
            gW[enc[2 * l + 1]], gB[enc[2 * l + 1]] = conv_wgrad(As[l], g)
            ga = q(conv_data(g, Wq[enc[2 * l + 1]])) * (As[l] > 0)
            gW[enc[2 * l]], gB[enc[2 * l]] = conv_wgrad(ins[l], ga)
            if l > 0:
                gin = q(conv_data(ga, Wq[enc[2 * l]]))
        flat = np.concatenate([np.concatenate([w.ravel(), b]) for w, b in zip(gW, gB)] +
                              [ghw, [ghb]])
        return loss, flat, y
