"""SpMV row schedule (host logic of libxamg_hip.so, no GPU): fold detection on
structured and unstructured matrices, and the tile schedule covers every row exactly
once with a grid the kernels accept (multiple of 8 workgroups, at most 1024 for
launches that write dot partials)."""
import numpy as np
import pytest
import scipy.sparse as sp


def lap1(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])


def csr_arrays(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.uint32), A.indices.astype(np.uint32)


def test_fold_cube(X):
    for N in (8, 16, 32):
        row, col, _ = X.cube(N)
        expect = N * N if N * N >= 256 else 0
        assert X.csr_fold(row, col) == expect


def test_fold_27_point():
    import pyxamg as X
    N = 16
    one = sp.diags([np.ones(N - 1), np.ones(N), np.ones(N - 1)], [-1, 0, 1])
    A = sp.kron(sp.kron(one, one), one)  # 27-point pattern
    row, col = csr_arrays(A)
    assert X.csr_fold(row, col) == N * N + N + 1


def test_fold_2d_and_none():
    import pyxamg as X
    N = 300
    A = sp.kronsum(lap1(N), lap1(N))  # 5-point
    row, col = csr_arrays(A)
    assert X.csr_fold(row, col) == N
    row, col = csr_arrays(lap1(5000))  # tridiagonal: largest offset 1 < 256
    assert X.csr_fold(row, col) == 0
    rng = np.random.default_rng(0)
    R = sp.random(4000, 4000, density=0.002, random_state=rng) + sp.eye(4000)
    row, col = csr_arrays(R)
    assert X.csr_fold(row, col) == 0
    # rectangular matrices are never folded
    row, col = csr_arrays(sp.random(300, 600, density=0.05, random_state=rng))
    assert X.csr_fold(row, col, ncols=600) == 0


CASES = [
    # nrows, fold, nv, prec, dots
    (128 ** 3, 128 ** 2, 32, 0, False),
    (128 ** 3, 128 ** 2, 32, 0, True),
    (128 ** 3, 128 ** 2, 8, 0, False),
    (128 ** 3, 128 ** 2, 16, 1, False),
    (128 ** 3, 128 ** 2 + 129, 32, 0, False),   # 27-point fold, not a tile multiple
    (100 ** 3, 100 ** 2, 32, 0, False),
    (100 ** 3, 100 ** 2, 64, 0, True),
    (1000 * 777, 1000, 64, 0, False),            # 2-D grid, ragged last fold row
    (123457, 5003, 64, 0, False),
    (123457, 5003, 1, 0, False),                 # nv=1: plain order (band fits L2)
    (5000, 0, 32, 0, False),
    (37, 0, 4, 1, True),
    (1, 0, 64, 0, False),
]


@pytest.mark.parametrize("n,fold,nv,prec,dots", CASES)
def test_schedule_covers_rows_once(X, n, fold, nv, prec, dots):
    grid, tpw, rows = X.spmv_schedule(n, 7 * n, fold, nv, prec, dots)
    assert grid % 8 == 0 and grid >= 8
    assert grid <= (1024 if dots else 16384)
    assert len(rows) <= grid * tpw
    b, e = rows[:, 0], rows[:, 1]
    assert (b <= e).all() and (b >= 0).all() and (e <= n).all()
    live = rows[e > b]
    order = np.argsort(live[:, 0])
    s = live[order]
    assert s[0, 0] == 0 and s[-1, 1] == n
    assert (s[1:, 0] == s[:-1, 1]).all()  # contiguous, no overlap, no gap


def test_schedule_folds_when_band_exceeds_l2(X):
    n, F = 128 ** 3, 128 ** 2
    grid, tpw, rows = X.spmv_schedule(n, 7 * n, F, 32, 0, False)
    # folded: consecutive tiles step by the fold within a strip, and each XCD's eighth
    # of the schedule stays inside one column strip of F/8 rows
    nt = len(rows)
    per_xcd = nt // 8
    for q in range(8):
        r = rows[q * per_xcd:(q + 1) * per_xcd]
        cols = np.unique(r[:, 0] % F // (F // 8))
        assert len(cols) == 1
    # a small band stays in plain row order
    grid, tpw, rows = X.spmv_schedule(n, 7 * n, F, 1, 0, False)
    assert (np.diff(rows[:, 0]) > 0).all()
    X.set_option("spmv_fold_enable", 0)
    try:
        _, _, rows = X.spmv_schedule(n, 7 * n, F, 32, 0, False)
        assert (np.diff(rows[:, 0]) > 0).all()
    finally:
        X.set_option("spmv_fold_enable", 1)
    with pytest.raises(X.XamgError):
        X.set_option("no_such_knob", 1)
