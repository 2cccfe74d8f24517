"""Driver for counter profiles of the finest-level SpMM: builds the 128^3 cube operator,
runs `reps` launches of y = A x for nv right-hand sides, and `reps` launches of the
library copy kernel over 8*n*nv bytes (the calibration kernel of known traffic).

--structure: instead, times (HIP events, 20 launches) the same SpMM kernel on the same
n rows and nv with couplings of growing reach: the diagonal only, x neighbours within grid
lines (3-point), x and y neighbours (5-point), and the cube's 7-point operator; prints one
JSON line per operator with microseconds per launch and the algorithmic GB/s."""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pyxamg as X  # noqa: E402

ap = argparse.ArgumentParser()
ap.add_argument("--nx", type=int, default=128)
ap.add_argument("--nv", type=int, default=32)
ap.add_argument("--reps", type=int, default=5)
ap.add_argument("--index-mode", type=int, default=0)
ap.add_argument("--structure", action="store_true")
a = ap.parse_args()
X.init(0)
n = a.nx ** 3


def banded(nx, offsets):
    """rows of the nx^3 grid coupled to the grid neighbours given as (dy, dx) offsets"""
    i = np.arange(nx ** 3, dtype=np.int64)
    cx, cy = i % nx, (i // nx) % nx
    cols = []
    for dy, dx in offsets:
        ok = (cx + dx >= 0) & (cx + dx < nx) & (cy + dy >= 0) & (cy + dy < nx)
        cols.append(np.where(ok, i + dy * nx + dx, -1))
    c = np.stack(cols, 1)
    rp = np.zeros(nx ** 3 + 1, np.uint32)
    rp[1:] = np.cumsum((c >= 0).sum(1))
    cc = c[c >= 0].astype(np.uint32)
    return rp, cc, np.ones(len(cc))


if a.structure:
    import json
    x = X.DeviceArray.from_host(np.random.default_rng(0).standard_normal(n * a.nv))
    y = X.DeviceArray(n * a.nv)
    ops = [("diagonal", banded(a.nx, [(0, 0)])),
           ("3-point x", banded(a.nx, [(0, -1), (0, 0), (0, 1)])),
           ("5-point xy", banded(a.nx, [(-1, 0), (0, -1), (0, 0), (0, 1), (1, 0)])),
           ("7-point cube", X.cube(a.nx))]
    for name, (rp, cc, vv) in ops:
        M = X.Matrix(rp, cc, vv, prec=0, index_mode=a.index_mode)
        ms = M.spmv_timed(a.nv, x, y, 20)
        alg = len(cc) * 10 + (n + 1) * 4 + 2 * n * a.nv * 8
        print(json.dumps(dict(operator=name, nnz=len(cc), us=round(ms * 1e3, 1),
                              alg_gbs=round(alg / ms / 1e6, 1))), flush=True)
    sys.exit(0)
row, col, val = X.cube(a.nx)
A = X.Matrix(row, col, val, prec=0, index_mode=a.index_mode)
x = X.DeviceArray.from_host(np.random.default_rng(0).standard_normal(n * a.nv))
y = X.DeviceArray(n * a.nv)
for _ in range(a.reps):
    A.spmv(a.nv, x, y)
for _ in range(a.reps):
    X.check(X.hip().xh_copy_device(y.ptr, x.ptr, n * a.nv), "xh_copy_device")
print("info", A.info(), "copy bytes read", 8 * n * a.nv, flush=True)
