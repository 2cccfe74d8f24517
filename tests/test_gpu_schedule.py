"""Folded (strip) row schedule of the SpMM on the GPU: the per-row arithmetic is
unchanged, so results stay bit-identical to the oracle whatever the tile order.
The fold threshold is lowered so that small operators take the folded path."""
import numpy as np
import pytest
import scipy.sparse as sp

from test_gpu_solver import levels_for, rel_diff

pytestmark = pytest.mark.gpu


@pytest.fixture
def forced_fold(gpu):
    X = gpu
    X.set_option("spmv_fold_min_bytes", 0)
    X.set_option("spmv_strip_budget", 64 << 10)  # many narrow strips
    try:
        yield X
    finally:
        X.set_option("spmv_fold_min_bytes", 2 << 20)
        X.set_option("spmv_strip_budget", 2 << 20)


@pytest.mark.parametrize("nv", [1, 8, 32, 64])
@pytest.mark.parametrize("index_mode", [1, 2])
def test_folded_spmv_bit_exact(forced_fold, O, nv, index_mode):
    X = forced_fold
    row, col, val = X.cube(24, 20, 18)
    n = len(row) - 1
    rng = np.random.default_rng(100 + nv)
    vals = rng.standard_normal(len(val))
    A = X.Matrix(row, col, vals, prec=0, index_mode=index_mode)
    assert A.fold() == 24 * 20
    g, tpw, tiles = X.spmv_schedule(n, len(col), 24 * 20, nv)
    assert len(tiles) > 8 and not (np.diff(tiles[:, 0]) > 0).all()  # really reordered
    x = rng.standard_normal(n * nv)
    dx, dy = X.DeviceArray.from_host(x), X.DeviceArray(n * nv)
    A.spmv(nv, dx, dy)
    assert np.array_equal(dy.to_host(), O.spmv(row, col, vals, x, nv))
    y0 = rng.standard_normal(n * nv)
    dy.upload(y0)
    A.spmv(nv, dx, dy, accumulate=True)
    assert np.array_equal(dy.to_host(), O.spmv(row, col, vals, x, nv, y=y0))


def test_folded_27_point_fp32(forced_fold, O):
    X = forced_fold
    N, nv = 18, 16
    one = sp.diags([np.ones(N - 1), np.ones(N), np.ones(N - 1)], [-1, 0, 1])
    A = sp.csr_matrix(sp.kron(sp.kron(one, one), one))
    A.sort_indices()
    row, col = A.indptr.astype(np.uint32), A.indices.astype(np.uint32)
    rng = np.random.default_rng(4)
    val = rng.standard_normal(len(col))
    M = X.Matrix(row, col, val, prec=1, index_mode=1)
    assert M.fold() == N * N + N + 1
    n = N ** 3
    x = rng.standard_normal(n * nv).astype(np.float32)
    dx, dy = X.DeviceArray.from_host(x), X.DeviceArray(n * nv, np.float32)
    M.spmv(nv, dx, dy)
    assert np.array_equal(dy.to_host(), O.spmv(row, col, val, x, nv, f32=True))


def test_fold_override(forced_fold, O):
    X = forced_fold
    row, col, val = X.cube(32)
    n, nv = len(row) - 1, 32
    A = X.Matrix(row, col, val)
    assert A.fold() == 1024
    x = X.rhs_uniform(2, n, nv)
    dx, dy = X.DeviceArray.from_host(x), X.DeviceArray(n * nv)
    ref = O.spmv(row, col, val, x, nv)
    for f in (0, 1024, 1000):
        assert A.fold(set=f) == f
        A.spmv(nv, dx, dy)
        assert np.array_equal(dy.to_host(), ref)
    with pytest.raises(X.XamgError):
        A.fold(set=n)


def test_folded_solver_and_preconditioner(forced_fold, O):
    X = forced_fold
    N, nv = 24, 8
    levels = levels_for(X, N)
    b = X.rhs_uniform(12, N ** 3, nv)
    prm = X.default_params(method="MergedPBiCGStab", smoother="Chebyshev", poly_order=4)
    S = X.Solver(X.Hier(levels), prm, nv)
    db, dx = X.DeviceArray.from_host(b), X.DeviceArray(N ** 3 * nv)
    S.precond_apply(db, dx)
    assert np.array_equal(dx.to_host(), O.mg_apply(O.Hier(levels), O.params_from(prm), nv, b))
    x = S.solve(b)
    it = S.stats()[0]
    ref = O.solve(O.Hier(levels), O.params_from(prm), nv, b)
    assert abs(it - ref["iters"]) <= 1
    assert rel_diff(x, ref["x"], nv) <= 1e-6
