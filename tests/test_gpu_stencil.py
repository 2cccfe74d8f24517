"""The 7-point stencil SpMV (register window along grid lines, option spmv_st7_enable)
against the row-parallel kernel.  It adds the same products in the same order, so every
launch without dot products gives bit-identical results: y = A x and y += A x for every
nv it takes, and a whole preconditioner application (Chebyshev and Jacobi smoothing,
residuals, restriction, prolongation) in fp64, with the reference's reduced-precision
hierarchy and with the fp32 preconditioner.  Full solves (whose dot partials are summed
per workgroup of another launch shape) meet the solver bar against the oracle.  Matrices
that are not a complete 7-point stencil in the generators' order get no stencil layout."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def st7(gpu):
    gpu.set_option("spmv_st7_enable", 1)
    yield gpu
    gpu.set_option("spmv_st7_enable", 0)
    gpu.set_option("spmv_st7_waves", 3072)
    gpu.set_option("spmv_st7_min_seg", 32)


def _mats(X):
    row, col, val = X.cube(13, 11, 9)
    yield "cube13x11x9", row, col, val
    r, c, v, _ = X.varcoef7(17, 6, 5, sigma=1.0)
    yield "varcoef7_17x6x5", r, c, v
    row, col, val = X.cube(5, 3, 2)
    yield "cube5x3x2", row, col, val


@pytest.mark.parametrize("prec", [0, 1])
def test_st7_spmv_bit_identical(st7, prec):
    X = st7
    dt = np.float64 if prec == 0 else np.float32
    for name, row, col, val in _mats(X):
        n = len(row) - 1
        A = X.Matrix(row, col, val, prec=prec)
        assert "st7" in A.layouts(), name
        for nv in (8, 16, 32, 64):
            # several launch shapes: many short segments, one wave per line block
            for waves, seg in ((3072, 32), (64, 1), (8, 4)):
                print(f"case {name} prec {prec} nv {nv} waves {waves} seg {seg}", flush=True)
                X.set_option("spmv_st7_waves", waves)
                X.set_option("spmv_st7_min_seg", seg)
                x = X.DeviceArray.from_host(
                    np.random.default_rng(nv).standard_normal(n * nv).astype(dt))
                out = []
                for on in (0, 1):
                    X.set_option("spmv_st7_enable", on)
                    y = X.DeviceArray(n * nv, dt)
                    A.spmv(nv, x, y)
                    y2 = X.DeviceArray.from_host(np.linspace(-1, 1, n * nv).astype(dt))
                    A.spmv(nv, x, y2, accumulate=True)
                    out.append((y.to_host(), y2.to_host()))
                X.set_option("spmv_st7_enable", 1)
                assert np.array_equal(out[0][0], out[1][0]), (name, nv, waves)
                assert np.array_equal(out[0][1], out[1][1]), (name, nv, waves)


def test_st7_layout_only_for_complete_stencils(gpu):
    X = gpu
    row, col, val = X.cube(8, 6, 5)
    assert "st7" in X.Matrix(row, col, val).layouts()
    # 27-point operator: grid view, no stencil layout
    r, c, v, _ = X.q1fem27(6, 5, 4)
    lay = X.Matrix(r, c, v).layouts()
    assert "grid" in lay and "st7" not in lay
    # one neighbour entry dropped: no stencil layout
    n = len(row) - 1
    keep = np.ones(len(col), bool)
    keep[row[n // 2] + 2] = False
    r2 = np.zeros(n + 1, np.int64)
    for i in range(n):
        r2[i + 1] = r2[i] + int(keep[row[i]:row[i + 1]].sum())
    lay = X.Matrix(r2.astype(np.uint32), col[keep], val[keep]).layouts()
    assert "grid" in lay and "st7" not in lay
    # neighbours stored in another order (+x before -x): no stencil layout
    c3, v3 = col.copy(), val.copy()
    i = n // 3
    s0, s1 = int(row[i]), int(row[i + 1])
    c3[s0 + 1:s1], v3[s0 + 1:s1] = c3[s0 + 1:s1][::-1].copy(), v3[s0 + 1:s1][::-1].copy()
    assert "st7" not in X.Matrix(row, c3, v3).layouts()


@pytest.mark.parametrize("mg_prec", [0, 1, 2])
@pytest.mark.parametrize("smoother,order", [("Chebyshev", 4), ("Chebyshev", 1), ("Jacobi", 2)])
def test_st7_preconditioner_bit_identical(st7, mg_prec, smoother, order):
    X = st7
    N, nv = 19, 8
    row, col, val = X.cube(N, 13, 11)
    n = len(row) - 1
    levels = X.Amg(row, col, val, agg_num_levels=1, coarse_size=100).export()
    prm = X.default_params(method="MergedPBiCGStab", precond="MultiGrid", smoother=smoother,
                           poly_order=order, sweeps=1 if smoother != "Jacobi" else order,
                           mg_prec=mg_prec, coarse_fp64=1 if mg_prec == 2 else 0)
    b = X.DeviceArray.from_host(X.rhs_uniform(3, n, nv))
    out = []
    for on in (0, 1):
        X.set_option("spmv_st7_enable", on)
        S = X.Solver(X.Hier(levels), prm, nv)
        x = X.DeviceArray(n * nv)
        S.precond_apply(b, x)
        out.append(x.to_host())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("method,nv", [("MergedPBiCGStab", 32), ("PBiCGStab", 8),
                                       ("RBiCGStab", 16), ("PCG", 64)])
def test_st7_solver_matches_oracle(st7, O, method, nv):
    X = st7
    row, col, val, _ = X.varcoef7(21, 18, 16, sigma=1.0, seed=3)
    levels = X.Amg(row, col, val, agg_num_levels=1).export()
    prm = X.default_params(method=method, smoother="Chebyshev", poly_order=4)
    b = X.rhs_uniform(1, len(row) - 1, nv)
    S = X.Solver(X.Hier(levels), prm, nv)
    x = S.solve(b)
    it, conv, _, _ = S.stats()
    ref = O.solve(O.Hier(levels), O.params_from(prm), nv, b)
    assert conv.all() and ref["conv"].all() and abs(it - ref["iters"]) <= 1
    a, r = x.reshape(-1, nv), ref["x"].reshape(-1, nv)
    assert (np.abs(a - r).max(0) / np.abs(r).max(0)).max() <= 1e-6
