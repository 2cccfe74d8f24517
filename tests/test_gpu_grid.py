"""The structured-grid SpMV (LDS-window line sweep, option spmv_grid_enable) against the
row-parallel kernel: the same products in the same order, so results are bit-identical;
and full solves with the grid path on against the oracle (the dot partials are summed in
another order, so the solver bar applies)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def grid_on(gpu):
    gpu.set_option("spmv_grid_enable", 1)
    yield gpu
    gpu.set_option("spmv_grid_enable", 2)  # the default: by row length


def _mats(X):
    row, col, val = X.cube(13, 11, 9)
    yield "cube13x11x9", row, col, val
    r, c, v, _ = X.q1fem27(10, 9, 7)
    yield "q1fem27_10x9x7", r, c, v
    r, c, v, _ = X.varcoef7(17, 6, 5, sigma=1.0)
    yield "varcoef7_17x6x5", r, c, v


@pytest.mark.parametrize("prec", [0, 1])
def test_grid_spmv_bit_identical(grid_on, prec):
    X = grid_on
    for name, row, col, val in _mats(X):
        n = len(row) - 1
        A = X.Matrix(row, col, val, prec=prec)
        assert A.grid() is not None, name
        dt = np.float64 if prec == 0 else np.float32
        for nv in (1, 2, 4, 8, 16, 32, 64):
            print(f"case {name} prec {prec} nv {nv}", flush=True)
            x = X.DeviceArray.from_host(np.random.default_rng(nv).standard_normal(n * nv).astype(dt))
            out = []
            for ge in (0, 1):
                X.set_option("spmv_grid_enable", ge)
                y = X.DeviceArray(n * nv, dt)
                A.spmv(nv, x, y)
                y2 = X.DeviceArray.from_host(np.linspace(-1, 1, n * nv).astype(dt))
                A.spmv(nv, x, y2, accumulate=True)
                out.append((y.to_host(), y2.to_host()))
            X.set_option("spmv_grid_enable", 1)
            assert np.array_equal(out[0][0], out[1][0]), (name, nv)
            assert np.array_equal(out[0][1], out[1][1]), (name, nv)


def test_grid_path_by_row_length(gpu):
    # the default takes the grid path for the 27-point operator (27 entries per interior
    # row) and the row-parallel kernel for the 7-point cube: the launches of a solve show it
    X = gpu
    X.set_option("spmv_grid_enable", 2)
    nv = 8
    for name, (row, col, val), want in (("q1fem27", X.q1fem27(12, 12, 12)[:3], True),
                                        ("cube", X.cube(12), False)):
        levels = X.Amg(row, col, val).export()
        prm = X.default_params(method="MergedPBiCGStab", smoother="Chebyshev", poly_order=2)
        S = X.Solver(X.Hier(levels), prm, nv)
        n = len(row) - 1
        db = X.DeviceArray.from_host(X.rhs_uniform(3, n, nv))
        dx = X.DeviceArray(n * nv)
        names = {k["kernel"] for k in S.profile(db, dx, 1)}
        assert any(k.startswith("k_spmv_grid<") for k in names) == want, (name, names)


def test_no_grid_view_for_partial_grids(gpu):
    X = gpu
    row, col, val = X.cube(8, 6, 5)
    n = len(row) - 1
    r2, c2, v2 = [0], [], []
    for i in range(n):
        cs, vs = list(col[row[i]:row[i + 1]]), list(val[row[i]:row[i + 1]])
        if i % 7 == 0:
            cs.append((i + n // 2) % n)
            vs.append(0.25)
        c2 += cs
        v2 += vs
        r2.append(len(c2))
    A = X.Matrix(np.array(r2, np.uint32), np.array(c2, np.uint32), np.array(v2))
    assert A.grid() is None
    assert X.Matrix(row, col, val).grid() == (8, 6, 5)


@pytest.mark.parametrize("method,smoother,nv", [("MergedPBiCGStab", "Chebyshev", 32),
                                                ("PBiCGStab", "HSGS", 8),
                                                ("RBiCGStab", "Jacobi", 4),
                                                ("PCG", "Chebyshev", 16)])
def test_solver_with_grid_path_matches_oracle(grid_on, O, method, smoother, nv):
    X = grid_on
    N = 20
    row, col, val = X.cube(N)
    levels = X.Amg(row, col, val).export()
    prm = X.default_params(method=method, smoother=smoother, poly_order=3,
                           sweeps=1 if smoother != "Jacobi" else 2)
    b = X.rhs_uniform(1, N ** 3, nv)
    S = X.Solver(X.Hier(levels), prm, nv)
    x = S.solve(b)
    it, conv, _, _ = S.stats()
    ref = O.solve(O.Hier(levels), O.params_from(prm), nv, b)
    assert conv.all() and ref["conv"].all() and abs(it - ref["iters"]) <= 1
    a, r = x.reshape(-1, nv), ref["x"].reshape(-1, nv)
    assert (np.abs(a - r).max(0) / np.abs(r).max(0)).max() <= 1e-6
