"""GPU kernel parity through the C ABI: SpMM and dense matvec bit-exact against the oracle
(same summation order, no FMA contraction on either side)."""
import numpy as np
import pytest
from conftest import random_csr

pytestmark = pytest.mark.gpu

NVS = [1, 2, 4, 8, 16, 32, 64]


@pytest.mark.parametrize("nv", NVS)
@pytest.mark.parametrize("index_mode", [1, 2])
def test_spmv_cube_bit_exact(gpu, O, nv, index_mode):
    X = gpu
    row, col, val = X.cube(24, 20, 18)
    n = len(row) - 1
    rng = np.random.default_rng(nv)
    vals = rng.standard_normal(len(val))
    A = X.Matrix(row, col, vals, prec=0, index_mode=index_mode)
    assert A.info()["index_bits"] == (32 if index_mode == 1 else 16)
    x = rng.standard_normal(n * nv)
    dx, dy = X.DeviceArray.from_host(x), X.DeviceArray(n * nv)
    A.spmv(nv, dx, dy)
    assert np.array_equal(dy.to_host(), O.spmv(row, col, vals, x, nv))
    y0 = rng.standard_normal(n * nv)
    dy.upload(y0)
    A.spmv(nv, dx, dy, accumulate=True)
    assert np.array_equal(dy.to_host(), O.spmv(row, col, vals, x, nv, y=y0))


@pytest.mark.parametrize("nv", NVS)
def test_spmv_fp32_bit_exact(gpu, O, nv):
    # rows up to 40 nonzeros: several chunks and batches for every lane-group width
    X = gpu
    rng = np.random.default_rng(11)
    row, col, val = random_csr(rng, 700, 700, 0.02, max_row=40)
    A = X.Matrix(row, col, val, prec=1, index_mode=1)
    x = rng.standard_normal(700 * nv).astype(np.float32)
    dx, dy = X.DeviceArray.from_host(x), X.DeviceArray(700 * nv, np.float32)
    A.spmv(nv, dx, dy)
    assert np.array_equal(dy.to_host(), O.spmv(row, col, val, x, nv, f32=True))


@pytest.mark.parametrize("nv", NVS)
def test_spmv_irregular_rows_bit_exact(gpu, O, nv):
    # empty rows, long rows (> 16 nonzeros: several cooperative chunks), rectangular
    X = gpu
    rng = np.random.default_rng(nv + 3)
    n, m = 1500, 900
    row, col, val = random_csr(rng, n, m, 0.03, diag=False)
    row = row.copy()
    lens = np.diff(row).astype(np.int64)
    lens[::7] = 0  # empty rows
    keep = np.concatenate([np.arange(row[i], row[i] + lens[i]) for i in range(n)]).astype(np.int64)
    col, val = col[keep], val[keep]
    row = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    assert lens.max() > 16
    A = X.Matrix(row, col, val, prec=0, index_mode=1, ncols=m)
    x = rng.standard_normal(m * nv)
    dx, dy = X.DeviceArray.from_host(x), X.DeviceArray(n * nv)
    A.spmv(nv, dx, dy)
    assert np.array_equal(dy.to_host(), O.spmv(row, col, val, x, nv, ncols=m))


@pytest.mark.parametrize("nv", [8, 32])
@pytest.mark.parametrize("prec", [0, 1])
def test_spmv_wide_batches_bit_exact(gpu, O, nv, prec):
    # few rows of many entries (the coarse levels of an aggressively coarsened hierarchy)
    # take the 16-wide batches: the same sums in the same order as the 8-wide kernel and,
    # in fp64, as the oracle; y = A x and y += A x, u16 and u32 columns, empty rows
    X = gpu
    rng = np.random.default_rng(nv * 7 + prec)
    n = 2000
    row, col, val = random_csr(rng, n, n, 0.04)
    lens = np.diff(row).astype(np.int64)
    lens[::11] = 0
    keep = np.concatenate([np.arange(row[i], row[i] + lens[i]) for i in range(n)]).astype(np.int64)
    col, val = col[keep], val[keep]
    row = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    assert len(col) >= 32 * n
    dt = np.float64 if prec == 0 else np.float32
    x = rng.standard_normal(n * nv)
    y0 = np.linspace(-1, 1, n * nv)
    try:
        for index_mode in (1, 2):
            A = X.Matrix(row, col, val, prec=prec, index_mode=index_mode)
            out = []
            for wide in (0, 1):
                X.set_option("spmv_wide_batches", wide)
                dx = X.DeviceArray.from_host(x.astype(dt))
                dy = X.DeviceArray(n * nv, dt)
                A.spmv(nv, dx, dy)
                dz = X.DeviceArray.from_host(y0.astype(dt))
                A.spmv(nv, dx, dz, accumulate=True)
                out.append((dy.to_host(), dz.to_host()))
            assert np.array_equal(out[0][0], out[1][0]), index_mode
            assert np.array_equal(out[0][1], out[1][1]), index_mode
            if prec == 0:
                assert np.array_equal(out[1][0], O.spmv(row, col, val, x, nv))
    finally:
        X.set_option("spmv_wide_batches", 1)


def test_wide_batches_in_the_coarse_levels(gpu):
    # the coarse levels of the aggressively coarsened 64^3 cube (levels 2 and 3: 2295 and
    # 587 rows of 34 and 45 entries) take the 16-wide SpMVs; the preconditioner is
    # bit-identical with them off
    X = gpu
    row, col, val = X.cube(64)
    levels = X.Amg(row, col, val, agg_num_levels=2, num_paths=2, coarse_size=500).export()
    n = len(row) - 1
    prm = X.default_params(method="MergedPBiCGStab", smoother="Chebyshev", poly_order=4)
    try:
        for nv in (8, 32):
            b = X.rhs_uniform(nv, n, nv)
            out, names = [], []
            for wide in (0, 1):
                X.set_option("spmv_wide_batches", wide)
                S = X.Solver(X.Hier(levels), prm, nv)
                db, dx = X.DeviceArray.from_host(b), X.DeviceArray(n * nv)
                S.precond_apply(db, dx)
                out.append(dx.to_host())
                dx.zero()
                names.append({(k["kernel"], k["level"]) for k in S.profile(db, dx, 1)})
            assert np.array_equal(out[0], out[1]), nv
            assert not any(k.endswith(",16>") for k, _ in names[0])
            assert any(k.endswith(",16>") and lv >= 2 for k, lv in names[1]), names[1]
            assert not any(k.endswith(",16>") and lv == 0 for k, lv in names[1])
    finally:
        X.set_option("spmv_wide_batches", 1)


def test_spmv_order_sensitive_row(gpu):
    # a row whose value depends on the summation order: (2^53 + 1) - 2^53 = 0 in the
    # reference's stored order
    X = gpu
    big = float(2 ** 53)
    row = np.array([0, 3], np.uint32)
    col = np.array([0, 1, 2], np.uint32)
    A = X.Matrix(row, col, np.array([big, 1.0, -big]), index_mode=1, ncols=3)
    dx, dy = X.DeviceArray.from_host(np.ones(3 * 4)), X.DeviceArray(4)
    A.spmv(4, dx, dy)
    assert np.array_equal(dy.to_host(), np.zeros(4))


def test_dense_matvec_bit_exact(gpu, O):
    import ctypes as C

    X = gpu
    rng = np.random.default_rng(5)
    n, nv = 97, 8
    D = rng.standard_normal((n, n))
    x = rng.standard_normal(n * nv)
    dD, dx, dy = X.DeviceArray.from_host(D.ravel()), X.DeviceArray.from_host(x), X.DeviceArray(n * nv)
    X.check(X.hip().xh_dense_matvec(n, nv, 0, dD.ptr, dx.ptr, dy.ptr), "xh_dense_matvec")
    y = np.zeros(n * nv)
    O.lib().xo_dense_matvec_d(n, n, D.ctypes.data_as(C.c_void_p), nv, x.ctypes.data_as(C.c_void_p),
                              y.ctypes.data_as(C.c_void_p))
    assert np.array_equal(dy.to_host(), y)


def test_spmv_full_size_cube_128(gpu, O):
    # the benchmark operator at full size: 2,097,152 rows, u16 panel columns, nv = 32
    X = gpu
    row, col, val = X.cube(128)
    n, nv = len(row) - 1, 32
    A = X.Matrix(row, col, val, prec=0, index_mode=0)
    info = A.info()
    assert info["index_bits"] == 16 and info["panel_rows"] == 32768
    x = X.rhs_uniform(3, n, nv)
    dx, dy = X.DeviceArray.from_host(x), X.DeviceArray(n * nv)
    A.spmv(nv, dx, dy)
    assert np.array_equal(dy.to_host(), O.spmv(row, col, val, x, nv))


def test_bad_arguments_are_rejected(gpu):
    X = gpu
    row, col, val = X.cube(4)
    A = X.Matrix(row, col, val)
    d = X.DeviceArray(64 * 3)
    assert X.hip().xh_spmv(A.m, 3, d.ptr, d.ptr, 0) < 0  # nv = 3 unsupported
    bad = col.copy()
    bad[0] = 10 ** 6
    with pytest.raises(X.XamgError):
        X.Matrix(row, bad, val)
    with pytest.raises(X.XamgError):  # a row spanning more than 65536 columns
        X.Matrix(np.array([0, 2], np.uint32), np.array([0, 100000], np.uint32), np.ones(2),
                 index_mode=2, ncols=100001)
