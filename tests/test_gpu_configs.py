"""Variable-coefficient configurations at test size: the 27-point trilinear-FE operator with
lognormal k (sigma = 2) under AMG + hybrid GS with the FP32 hierarchy, and the 7-point
finite-volume operator (sigma = 1) under BiCGStab preconditioned by a degree-8 Chebyshev
polynomial without AMG.  GPU against the oracle on identical matrices and hierarchies:
iteration counts within 1 and per-RHS relative solution difference <= 1e-6."""
import numpy as np
import pytest

from test_gpu_solver import rel_diff, run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mg_prec", [0, 2])
def test_q1fem27_lognormal_hsgs(gpu, O, mg_prec):
    X = gpu
    N, nv = 24, 8
    row, col, val, _ = X.q1fem27(N, sigma=2.0, seed=1)
    levels = X.Amg(row, col, val, agg_num_levels=2, num_paths=2, coarse_size=500).export()
    b = X.rhs_uniform(3, N ** 3, nv)
    prm = X.default_params(method="MergedPBiCGStab", smoother="HSGS", sweeps=1, max_iters=300,
                           rel_tol=1e-8, mg_prec=mg_prec)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["conv"].all() and r["conv"].all()
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel_diff(g["x"], r["x"], nv) <= 1e-6
    if mg_prec == 2:
        # the FP32-hierarchy solution against the FP64-hierarchy solve of the same system
        p64 = X.default_params(method="MergedPBiCGStab", smoother="HSGS", sweeps=1,
                               max_iters=300, rel_tol=1e-8, mg_prec=0)
        x64 = X.Solver(X.Hier(levels), p64, nv).solve(b)
        assert rel_diff(g["x"], x64, nv) <= 1e-6


@pytest.mark.parametrize("mg_prec", [0, 2])
def test_q1fem27_hsgs_preconditioner_bit_exact(gpu, O, mg_prec):
    # one V-cycle with hybrid GS on the 27-point operator, whose finest residuals take the
    # structured-grid path by default (the launches show it): bit-identical to the oracle's
    X = gpu
    N, nv = 24, 8
    row, col, val, _ = X.q1fem27(N, sigma=2.0, seed=1)
    levels = X.Amg(row, col, val, agg_num_levels=2, num_paths=2, coarse_size=500).export()
    prm = X.default_params(method="MergedPBiCGStab", smoother="HSGS", sweeps=1, mg_prec=mg_prec)
    S = X.Solver(X.Hier(levels), prm, nv)
    b = X.rhs_uniform(9, N ** 3, nv)
    db, dx = X.DeviceArray.from_host(b), X.DeviceArray(N ** 3 * nv)
    S.precond_apply(db, dx)
    assert np.array_equal(dx.to_host(), O.mg_apply(O.Hier(levels), O.params_from(prm), nv, b))
    dx.zero()
    names = {k["kernel"] for k in S.profile(db, dx, 1)}
    prec = "double" if mg_prec == 0 else "float"
    assert f"k_spmv_grid<{prec},{nv},OP_RESID>" in names, names


def test_varcoef7_chebyshev8_without_amg(gpu, O):
    X = gpu
    N, nv = 20, 4
    row, col, val, _ = X.varcoef7(N, sigma=1.0, seed=2)
    levels = X.Amg(row, col, val, max_levels=1).export()
    b = X.rhs_uniform(4, N ** 3, nv)
    prm = X.default_params(method="PBiCGStab", precond="Chebyshev", poly_order=8, max_iters=500,
                           rel_tol=1e-8)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["conv"].all() and r["conv"].all()
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel_diff(g["x"], r["x"], nv) <= 1e-6


def test_device_bump_rhs_matches_host_generator(gpu):
    X = gpu
    b_host = X.rhs_bumps(17, 24, 20, 16, 4)
    b_dev = X.rhs_bumps_device(17, 24, 20, 16, 4).to_host()
    assert np.allclose(b_dev, b_host, rtol=1e-12, atol=1e-14 * np.abs(b_host).max())


def test_q1fem27_full_size_fp32_hierarchy(gpu):
    """The 27-point configuration at its full size (96^3, 23.4 M nonzeros, nv = 32) under
    MergedPBiCGStab + AMG with hybrid GS, fp64 and fp32 hierarchy.  The oracle pins both at
    24^3 (above); here the checks are size-independent: every RHS converges, the true
    residual ||b - A x|| / ||b|| of each RHS (recomputed on the host) is at the tolerance,
    and the fp32-hierarchy solution agrees with the fp64 one to 1e-6 per RHS."""
    import time

    import scipy.sparse as sp

    X = gpu
    N, nv = 96, 32
    row, col, val, _ = X.q1fem27(N, sigma=2.0, seed=1)
    assert len(col) == (3 * N - 2) ** 3
    levels = X.Amg(row, col, val, agg_num_levels=2, num_paths=2, coarse_size=500).export()
    H = X.Hier(levels)
    b = X.rhs_uniform(5, N ** 3, nv)
    A = sp.csr_matrix((val, col.astype(np.int64), row.astype(np.int64)), shape=(N ** 3, N ** 3))
    bm = b.reshape(-1, nv)
    xs = {}
    for mg_prec in (0, 2):
        prm = X.default_params(method="MergedPBiCGStab", smoother="HSGS", sweeps=1,
                               max_iters=300, rel_tol=1e-8, abs_tol=0.0, mg_prec=mg_prec,
                               coarse_fp64=1)
        S = X.Solver(H, prm, nv)
        db, dx = X.DeviceArray.from_host(b), X.DeviceArray(N ** 3 * nv)
        S.solve_device(db, dx)  # warm-up (graph capture)
        dx.zero()
        X.hip().xh_synchronize()
        t = time.perf_counter()
        S.solve_device(db, dx)
        X.hip().xh_synchronize()
        ms = (time.perf_counter() - t) * 1e3
        it, conv, _, rel = S.stats()
        x = dx.to_host().reshape(-1, nv)
        true_rel = np.linalg.norm(bm - A @ x, axis=0) / np.linalg.norm(bm, axis=0)
        print(f"q1fem27 96^3 nv=32 mg_prec={mg_prec}: {it} iterations, {ms:.1f} ms "
              f"({ms / nv:.3f} ms/RHS), max true rel res {true_rel.max():.2e}")
        assert conv.all()
        assert true_rel.max() <= 2e-8
        xs[mg_prec] = x
    diff = rel_diff(xs[2], xs[0], nv)
    print(f"fp32 hierarchy vs fp64: max per-RHS rel diff {diff:.2e}")
    assert diff <= 1e-6
