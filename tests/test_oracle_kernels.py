"""Pins of the oracle's kernel-level functions against independent computations:
dense brute force, integer-exact arithmetic, closed forms of the cube operator,
numpy.linalg / numpy.polynomial."""
import numpy as np
import pytest
from conftest import random_csr, to_dense


@pytest.mark.parametrize("nv", [1, 3, 4, 8])
def test_spmv_integer_exact(O, nv):
    rng = np.random.default_rng(nv)
    row, col, val = random_csr(rng, 57, 57, 0.15, int_vals=True)
    x = rng.integers(-5, 6, size=57 * nv).astype(np.float64)
    D = to_dense(row, col, val)
    y = O.spmv(row, col, val, x, nv)
    assert np.array_equal(y.reshape(-1, nv), D @ x.reshape(-1, nv))
    y0 = rng.integers(-5, 6, size=57 * nv).astype(np.float64)
    y2 = O.spmv(row, col, val, x, nv, y=y0)
    assert np.array_equal(y2.reshape(-1, nv), y0.reshape(-1, nv) + D @ x.reshape(-1, nv))


def test_spmv_float_and_rectangular(O):
    rng = np.random.default_rng(7)
    row, col, val = random_csr(rng, 40, 23, 0.2, diag=False)
    x = rng.standard_normal(23 * 4)
    D = to_dense(row, col, val, 23)
    y = O.spmv(row, col, val, x, 4, ncols=23)
    np.testing.assert_allclose(y.reshape(-1, 4), D @ x.reshape(-1, 4), rtol=1e-13, atol=1e-13)
    yf = O.spmv(row, col, val, x.astype(np.float32), 4, ncols=23, f32=True)
    np.testing.assert_allclose(yf.reshape(-1, 4), D @ x.reshape(-1, 4), rtol=2e-5, atol=2e-5)


def test_spmv_sequential_order(O):
    # 2^53 + 1 - 2^53: the order of the row sum decides the result (0 or 1); the
    # reference sums in stored order
    big = float(2 ** 53)
    row = np.array([0, 3], np.uint32)
    col = np.array([0, 1, 2], np.uint32)
    val = np.array([big, 1.0, -big])
    y = O.spmv(row, col, val, np.ones(3), 1, ncols=3)
    assert y[0] == 0.0  # (2^53 + 1) rounds to 2^53 (ties to even)
    val2 = np.array([big, -big, 1.0])
    assert O.spmv(row, col, val2, np.ones(3), 1, ncols=3)[0] == 1.0


@pytest.mark.parametrize("dims", [(5, 5, 5), (8, 3, 6), (1, 7, 4), (16, 16, 16)])
def test_cube_generator_closed_forms(O, dims):
    nx, ny, nz = dims
    row, col, val = O.cube(nx, ny, nz)
    n = nx * ny * nz
    # nnz = 7N minus one per missing neighbour: 2 N / d per dimension of size d
    nnz = 7 * n - sum(2 * n // d for d in dims)
    assert len(col) == nnz
    # diagonal first, then ascending columns
    for i in range(0, n, max(1, n // 50)):
        c = col[row[i]:row[i + 1]]
        assert c[0] == i and np.all(np.diff(c[1:].astype(np.int64)) > 0)
    # eigenpairs of the ghost-cell Dirichlet operator: v = prod sin((i+1/2) pi m / n)
    rng = np.random.default_rng(0)
    for _ in range(3):
        m = [rng.integers(1, d + 1) for d in dims]
        q = [float(d) ** 2 for d in dims]
        ii = np.arange(n) % nx
        jj = (np.arange(n) // nx) % ny
        kk = np.arange(n) // (nx * ny)
        v = np.ones(n)
        lam = 0.0
        for a, (idx, d, mm, qq) in enumerate(zip((ii, jj, kk), dims, m, q)):
            if d == 1:
                continue
            v *= np.sin((idx + 0.5) * np.pi * mm / d)
            lam += 4 * qq * np.sin(np.pi * mm / (2 * d)) ** 2
        Av = O.spmv(row, col, val, v, 1)
        np.testing.assert_allclose(Av, lam * v, rtol=1e-9, atol=1e-9 * lam)


@pytest.mark.parametrize("n", [1, 5, 33])
def test_gauss_jordan_inverse(O, n):
    import ctypes as C

    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    inv = np.zeros((n, n))
    O.lib().xo_inverse_d(n, A.ctypes.data_as(C.c_void_p), inv.ctypes.data_as(C.c_void_p))
    np.testing.assert_allclose(inv, np.linalg.inv(A), rtol=1e-12, atol=1e-12)
    Af = A.astype(np.float32)
    invf = np.zeros((n, n), np.float32)
    O.lib().xo_inverse_f(n, Af.ctypes.data_as(C.c_void_p), invf.ctypes.data_as(C.c_void_p))
    np.testing.assert_allclose(invf, np.linalg.inv(A), rtol=1e-4, atol=1e-5)


def test_dense_matvec_integer_exact(O):
    import ctypes as C

    rng = np.random.default_rng(3)
    n, nv = 17, 4
    D = rng.integers(-4, 5, size=(n, n)).astype(np.float64)
    x = rng.integers(-4, 5, size=n * nv).astype(np.float64)
    y = np.zeros(n * nv)
    O.lib().xo_dense_matvec_d(n, n, D.ctypes.data_as(C.c_void_p), nv,
                              x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p))
    assert np.array_equal(y.reshape(-1, nv), D @ x.reshape(-1, nv))


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6, 8])
def test_chebyshev_coefficients(O, order):
    # p(t) = (1 - T_k((th - t)/de) / T_k(th/de)) / t with the reference's bounds
    # ub = 1.1 max, lb = min + fraction (ub - min)
    from numpy.polynomial import Chebyshev, Polynomial

    mn, mx, fr = 0.05, 1.9, 0.3
    ub = 1.1 * mx
    lb = (ub - mn) * fr + mn
    th, de = (ub + lb) / 2, (ub - lb) / 2
    Tk = Chebyshev.basis(order).convert(kind=Polynomial)
    s = Polynomial([th / de, -1.0 / de])  # s(t) = (th - t)/de
    r = Tk(s) / Tk(th / de)
    p = (Polynomial([1.0]) - r)
    coef = p.coef[1:]  # divide by t: the constant term of 1 - r is 0
    assert abs(p.coef[0]) < 1e-12
    c = O.cheby_coeffs(mn, mx, order, fr)
    np.testing.assert_allclose(c, coef[:order], rtol=1e-10, atol=1e-12)


def test_split_range_and_core_ranges(O):
    import ctypes as C

    # split_range rounds n / P up to a multiple of 8 and always adds (misc.h:46-55)
    assert O.lib().xo_split_range(64, 1) == 72
    assert O.lib().xo_split_range(100, 3) == 40
    assert O.lib().xo_split_range(7, 1) == 8
    off = np.zeros(5, np.uint64)
    O.lib().xo_core_ranges(100, 4, off.ctypes.data_as(C.c_void_p))
    # core_size = 25 + 7 = 32: blocks [0,32) [32,64) [64,96) [96,100)
    assert off.tolist() == [0, 32, 64, 96, 100]


def test_define_int_type_thresholds(O):
    f = O.lib().xo_define_int_type
    assert [f(255), f(256), f(65535), f(65536), f(2 ** 32 - 1), f(2 ** 32)] == [0, 1, 1, 2, 2, 3]


def test_compress_round_trip(O):
    import ctypes as C

    rng = np.random.default_rng(5)
    n, ncols = 30, 500
    row, col, val = random_csr(rng, n, ncols, 0.01, diag=False)
    # make some rows empty
    row = row.copy()
    A = O.csr(row, col, val, ncols)
    r_out = np.zeros(n + 1, np.uint32)
    c_out = np.zeros(len(col), np.uint32)
    r_ind = np.zeros(n, np.uint32)
    c_ind = np.zeros(ncols, np.uint32)
    nr, nc = C.c_int64(), C.c_int64()
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    O.lib().xo_compress(C.byref(A), ncols, p(r_out), p(c_out), p(r_ind), p(c_ind), C.byref(nr),
                        C.byref(nc))
    nr, nc = nr.value, nc.value
    nonempty = [i for i in range(n) if row[i + 1] > row[i]]
    assert nr == len(nonempty) and r_ind[:nr].tolist() == nonempty
    used = sorted(set(col.tolist()))
    assert nc == len(used) and c_ind[:nc].tolist() == used  # ascending by construction
    # decode: col_ind[col[j]] == original column; rows map back through row_ind
    dec = c_ind[c_out]
    assert np.array_equal(dec, col)
    for q in range(nr):
        i = r_ind[q]
        assert np.array_equal(dec[r_out[q]:r_out[q + 1]], col[row[i]:row[i + 1]])
    assert O.lib().xo_define_int_type(nc) == (0 if nc <= 255 else 1)
