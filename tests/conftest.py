import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "softx-d-21-00014_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_built():
    if not os.path.exists(os.path.join(ORACLE, "build", "libxamg_oracle.so")):
        subprocess.check_call(["make", "-s", "-C", ORACLE])
    libs = [os.path.join(PKG, "lib", n) for n in ("libxamg_hip.so", "libxamg.so")]
    if not all(os.path.exists(p) for p in libs):
        subprocess.check_call(["make", "-s", "-j8", "-C", PKG])


_ensure_built()


@pytest.fixture(scope="session")
def X():
    import pyxamg

    return pyxamg


@pytest.fixture(scope="session")
def O():
    import pyoracle

    return pyoracle


@pytest.fixture(scope="session")
def gpu(X):
    X.init(0)
    return X


def random_csr(rng, n, m, density, int_vals=False, diag=True, sort=True, max_row=None):
    """random CSR (u32 row/col, f64 val); optional diagonal first then ascending columns"""
    rows, cols, vals = [0], [], []
    for i in range(n):
        k = rng.binomial(m, density)
        if max_row is not None:
            k = min(k, max_row)
        cs = set(rng.choice(m, size=k, replace=False).tolist()) if k else set()
        if diag and i < m:
            cs.add(i)
        cs = sorted(cs)
        if sort and diag and i in cs:
            cs.remove(i)
            cs = [i] + cs
        for c in cs:
            cols.append(c)
            if int_vals:
                v = float(rng.integers(-3, 4))
                if c == i:
                    v = float(rng.integers(8, 12))
            else:
                v = rng.standard_normal()
                if c == i:
                    v = 4.0 + abs(v) * len(cs)
            vals.append(v)
        rows.append(len(cols))
    return (np.array(rows, np.uint32), np.array(cols, np.uint32), np.array(vals, np.float64))


def to_dense(row, col, val, m=None):
    n = len(row) - 1
    m = n if m is None else m
    D = np.zeros((n, m))
    for i in range(n):
        for j in range(row[i], row[i + 1]):
            D[i, col[j]] += val[j]
    return D
