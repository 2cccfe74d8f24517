"""Driver for counter profiles of the finest-level SpMM: builds the 128^3 cube operator,
runs `reps` launches of y = A x for nv right-hand sides, and `reps` launches of the
library copy kernel over 8*n*nv bytes (the calibration kernel of known traffic)."""
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
a = ap.parse_args()
X.init(0)
n = a.nx ** 3
row, col, val = X.cube(a.nx)
A = X.Matrix(row, col, val, prec=0, index_mode=a.index_mode)
x = X.DeviceArray.from_host(np.random.default_rng(0).standard_normal(n * a.nv))
y = X.DeviceArray(n * a.nv)
for _ in range(a.reps):
    A.spmv(a.nv, x, y)
for _ in range(a.reps):
    X.check(X.hip().xh_copy_device(y.ptr, x.ptr, n * a.nv), "xh_copy_device")
print("info", A.info(), "copy bytes read", 8 * n * a.nv, flush=True)
