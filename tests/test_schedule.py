"""SpMV row schedule (host logic of libxamg_hip.so, no GPU): fold detection on
structured and unstructured matrices, and the tile schedule covers every row exactly
once with a grid the kernels accept (multiple of 8 workgroups, at most 1024 for
launches that write dot partials); the structured-grid view (dimensions and per-entry
neighbour codes against a direct restatement) and its work items."""
import numpy as np
import pytest
import scipy.sparse as sp
from conftest import random_csr


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


# ---- structured-grid view (xh_csr_grid, xh_grid_schedule) ----

def _numpy_codes(row, col, nx, ny, nz):
    n = len(row) - 1
    F = nx * ny
    code = np.full(len(col), 255, np.uint8)
    for i in range(n):
        c, y, z = i % nx, (i // nx) % ny, i // F
        for j in range(row[i], row[i + 1]):
            dl = int(col[j]) - i
            for a in (-1, 0, 1):
                for b in (-1, 0, 1):
                    for d in (-1, 0, 1):
                        if dl == a * F + b * nx + d and 0 <= c + d < nx and 0 <= y + b < ny \
                                and 0 <= z + a < nz:
                            code[j] = 9 * (a + 1) + 3 * (b + 1) + (d + 1)
    return code


@pytest.mark.parametrize("gen,dims", [("cube", (10, 7, 5)), ("cube", (3, 3, 3)),
                                      ("q1fem27", (6, 5, 4)), ("varcoef7", (9, 4, 6))])
def test_grid_view_detection_and_codes(X, gen, dims):
    if gen == "cube":
        row, col, val = X.cube(*dims)
    elif gen == "q1fem27":
        row, col, val, _ = X.q1fem27(*dims)
    else:
        row, col, val, _ = X.varcoef7(*dims)
    d, code, coded = X.csr_grid(row, col)
    assert d == dims and coded == len(col)
    np.testing.assert_array_equal(code, _numpy_codes(row, col, *dims))


def test_grid_view_absent_for_unstructured_and_partial(X):
    rng = np.random.default_rng(0)
    row, col, val = random_csr(rng, 200, 200, 0.05)
    assert X.csr_grid(row, col)[0] is None
    # a grid operator with one off-grid coupling per row of the first plane: a view with
    # coded < nnz, which the upload does not attach
    row, col, val = X.cube(8, 6, 5)
    n = len(row) - 1
    r2, c2 = [0], []
    for i in range(n):
        cs = list(col[row[i]:row[i + 1]])
        if i < 48:
            cs.append((i + n // 2) % n)
        c2 += cs
        r2.append(len(c2))
    d, code, coded = X.csr_grid(np.array(r2, np.uint32), np.array(c2, np.uint32))
    assert d == (8, 6, 5) and coded == len(c2) - 48 and (code == 255).sum() == 48


@pytest.mark.parametrize("dims,nv,prec,dots", [((128, 128, 128), 32, 0, False),
                                               ((128, 128, 128), 32, 0, True),
                                               ((13, 11, 9), 8, 0, False),
                                               ((20, 7, 31), 1, 1, True),
                                               ((64, 40, 3), 64, 0, True)])
def test_grid_schedule_covers_every_row_once(X, dims, nv, prec, dots):
    nx, ny, nz = dims
    g, ipw, items = X.grid_schedule(nx, ny, nz, nv, prec, dots)
    assert g % 8 == 0 and g * ipw >= len(items)
    if dots:
        assert g <= 1024
    w = 16 // (8 if prec == 0 else 4)
    rpb = 256 // (nv // min(w, nv))
    tz = 16 if rpb >= 256 else 8 if rpb >= 64 else 4 if rpb >= 16 else 2 if rpb >= 4 else 1
    ty = rpb // tz
    hit = np.zeros((nz, ny, nx), np.int32)
    for y0, z0, c0, c1 in items:
        assert 0 <= c0 < c1 <= nx
        hit[z0:z0 + tz, y0:y0 + ty, c0:c1] += 1
    assert (hit == 1).all()
