"""Driver for kernel profiles of U-Net inference: a 4-level network (16/32/48/64 channels,
random initial weights) maps nv right-hand sides of an nx^3 grid (standard normal, seed 0)
to x0, `reps` times after one untimed inference; prints the time per RHS and a checksum
of x0.  --no-tiles runs every convolution through the per-tap gather kernel instead of
the LDS-staged tiles (same output)."""
import argparse
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pyxamg as X  # noqa: E402

ap = argparse.ArgumentParser()
ap.add_argument("--nx", type=int, default=128)
ap.add_argument("--nv", type=int, default=32)
ap.add_argument("--reps", type=int, default=2)
ap.add_argument("--no-tiles", action="store_true")
a = ap.parse_args()
X.init(0)
X.set_option("unet_conv_tile", 0 if a.no_tiles else 1)
net = X.UNet((16, 32, 48, 64), seed=1)
net.output_scale = 0.37
n, nv = a.nx, a.nv
b = X.DeviceArray.from_host(np.random.default_rng(0).standard_normal(n ** 3 * nv))
x = X.DeviceArray(n ** 3 * nv)
net.infer(n, n, n, nv, b, x)
X.hip().xh_synchronize()
t0 = time.perf_counter()
for _ in range(a.reps):
    net.infer(n, n, n, nv, b, x)
X.hip().xh_synchronize()
t1 = time.perf_counter()
print(f"inference {n}^3 nv={nv}: {(t1 - t0) / a.reps * 1e3 / nv:.4f} ms/RHS, "
      f"checksum {np.abs(x.to_host()).sum():.6e}", flush=True)
