"""GPU solve parity through the C ABI against the oracle on identical hierarchies.

Bar (from the task's correctness definition): iteration counts within +-1 and per-RHS
relative solution difference <= 1e-6 on fp64 paths.  The preconditioner application
contains no reductions, so it is compared bit for bit."""
import os

import numpy as np
import pytest

import yaml_params

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_cache = {}


def levels_for(X, N, **kw):
    key = (N, tuple(sorted(kw.items())))
    if key not in _cache:
        row, col, val = X.cube(N)
        _cache[key] = X.Amg(row, col, val, **kw).export()
    return _cache[key]


def rel_diff(a, b, nv):
    a, b = a.reshape(-1, nv), b.reshape(-1, nv)
    return (np.abs(a - b).max(0) / np.abs(b).max(0)).max()


def run_both(X, O, levels, prm, nv, b, x0=None, hist=False):
    S = X.Solver(X.Hier(levels), prm, nv)
    x = S.solve(b, x0)
    it, conv, a, r = S.stats()
    ref = O.solve(O.Hier(levels), O.params_from(prm), nv, b, x0,
                  hist_len=prm.solver.max_iters + 2 if hist else 0)
    out = dict(x=x, iters=it, conv=conv, abs_res=a, rel_res=r)
    if hist:
        out["hist"] = S.history()
    return out, ref


CASES = [
    # (name, N, nv, method, smoother, order/sweeps, extra params)
    ("C1_hsgs", 32, 4, "PBiCGStab", "HSGS", 1, {}),
    ("cheb4_merged", 32, 8, "MergedPBiCGStab", "Chebyshev", 4, {}),
    ("jacobi2", 24, 2, "PBiCGStab", "Jacobi", 2, {}),
    ("cheb2_nv1", 20, 1, "PBiCGStab", "Chebyshev", 2, {}),
    ("cheb3_nv64", 16, 64, "MergedPBiCGStab", "Chebyshev", 3, {}),
    ("wcycle", 20, 4, "PBiCGStab", "Chebyshev", 2, {"mg_cycle": 1}),
    ("fcycle", 20, 4, "PBiCGStab", "HSGS", 1, {"mg_cycle": 2}),
    ("hsgs_exact_gs", 16, 2, "PBiCGStab", "HSGS", 1, {"hsgs_block_rows": 0}),
    ("no_graph", 20, 4, "MergedPBiCGStab", "Chebyshev", 4, {"graph": 0}),
    ("u32", 20, 4, "MergedPBiCGStab", "Chebyshev", 4, {"index_mode": 1}),
    ("pcg_hsgs", 24, 4, "PCG", "HSGS", 1, {}),
    ("pcg_cheb3_nv32", 16, 32, "PCG", "Chebyshev", 3, {}),
    ("pcg_jacobi_nv1", 20, 1, "PCG", "Jacobi", 2, {}),
    ("rbicgstab_cheb2", 24, 8, "RBiCGStab", "Chebyshev", 2, {}),
    ("rbicgstab_hsgs_nograph", 20, 4, "RBiCGStab", "HSGS", 1, {"graph": 0}),
    ("merged_rbicgstab_jacobi", 20, 4, "MergedRBiCGStab", "Jacobi", 2, {}),
    ("merged_rbicgstab_nv64", 16, 64, "MergedRBiCGStab", "Chebyshev", 3, {}),
    ("pcg_fp32_precond", 20, 8, "PCG", "Chebyshev", 3, {"mg_prec": 2}),
    ("merged_rbicgstab_wcycle", 16, 4, "MergedRBiCGStab", "Chebyshev", 2, {"mg_cycle": 1}),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_solver_parity(gpu, O, case):
    X = gpu
    name, N, nv, method, smoother, k, extra = case
    levels = levels_for(X, N)
    prm = X.default_params(method=method, smoother=smoother, poly_order=k,
                           sweeps=1 if smoother in ("HSGS", "Chebyshev") else k)
    for kk, v in extra.items():
        setattr(prm, kk, v)
    b = X.rhs_uniform(0, N ** 3, nv)
    g, r = run_both(X, O, levels, prm, nv, b, hist=True)
    assert g["conv"].all() and r["conv"].all()
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel_diff(g["x"], r["x"], nv) <= 1e-6
    if g["iters"] == r["iters"]:
        # same iteration path: the per-check residual histories agree to rounding level
        # (the recurrence residuals theta - omega*phi and eta - omega*theta are differences
        # of nearly equal terms near convergence: rounding relative to the initial norm).
        # With an fp32 preconditioner the fp64 residual is rounded to fp32 before each
        # V-cycle, so a last-bit difference from the dot order can flip one fp32 rounding:
        # the histories then agree to the fp32 scale.
        fp32 = extra.get("mg_prec", 0) > 0
        np.testing.assert_allclose(g["hist"], r["hist"], rtol=1e-5 if fp32 else 1e-8,
                                   atol=(1e-8 if fp32 else 1e-11) * r["hist"][0].max())


@pytest.mark.parametrize("method", ["BiCGStab", "MergedBiCGStab"])
def test_unpreconditioned_parity(gpu, O, method):
    X = gpu
    N, nv = 12, 4
    row, col, val = X.cube(N)
    levels = [{"A": (N ** 3, N ** 3, row, col, val), "eig": (0.0, 2.0)}]
    prm = X.default_params(method=method, precond=None, max_iters=400, rel_tol=1e-8)
    b = X.rhs_uniform(2, N ** 3, nv)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["conv"].all() and abs(g["iters"] - r["iters"]) <= 1
    assert rel_diff(g["x"], r["x"], nv) <= 1e-6


@pytest.mark.parametrize("order", [4, 8])
def test_chebyshev_preconditioner_without_multigrid(gpu, O, order):
    # BiCGStab + Jacobi-scaled Chebyshev (order 8 is beyond the reference's 1-4)
    X = gpu
    N, nv = 16, 4
    row, col, val = X.cube(N)
    amg = X.Amg(row, col, val, max_levels=1)
    levels = amg.export()
    prm = X.default_params(method="PBiCGStab", precond="Chebyshev", smoother="Chebyshev",
                           poly_order=order, max_iters=300)
    prm.precond = X.role("Chebyshev", poly_order=order)
    b = X.rhs_uniform(4, N ** 3, nv)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["conv"].all() and abs(g["iters"] - r["iters"]) <= 1
    assert rel_diff(g["x"], r["x"], nv) <= 1e-6


@pytest.mark.parametrize("mg_prec", [1, 2])
def test_mixed_precision_parity(gpu, O, mg_prec):
    # 1: the reference's mg_reduced_precision (levels >= 1 and transfers in fp32);
    # 2: fp32 preconditioner including the finest level, fp64 coarsest solve
    X = gpu
    N, nv = 32, 4
    levels = levels_for(X, N)
    prm = X.default_params(method="MergedPBiCGStab", smoother="Chebyshev", poly_order=4,
                           mg_prec=mg_prec)
    b = X.rhs_uniform(0, N ** 3, nv)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["conv"].all() and abs(g["iters"] - r["iters"]) <= 1
    assert rel_diff(g["x"], r["x"], nv) <= 1e-6
    # error against the fp64 solve of the same system
    p64 = X.default_params(method="MergedPBiCGStab", smoother="Chebyshev", poly_order=4)
    x64 = X.Solver(X.Hier(levels), p64, nv).solve(b)
    assert rel_diff(g["x"], x64, nv) <= 1e-6


# Chebyshev order 1 post-smoothing updates a non-zero x from r = b - A x: the x update
# runs after the SpMV (an in-place update in the same launch raced with other rows' reads)
@pytest.mark.parametrize("smoother,k", [("Chebyshev", 4), ("Chebyshev", 1), ("HSGS", 1),
                                        ("Jacobi", 2)])
@pytest.mark.parametrize("mg_prec", [0, 1, 2])
def test_preconditioner_bit_exact(gpu, O, smoother, k, mg_prec):
    X = gpu
    N, nv = 20, 8
    levels = levels_for(X, N)
    prm = X.default_params(method="MergedPBiCGStab", smoother=smoother, poly_order=k,
                           sweeps=1 if smoother != "Jacobi" else k, mg_prec=mg_prec)
    S = X.Solver(X.Hier(levels), prm, nv)
    b = X.rhs_uniform(9, N ** 3, nv)
    db, dx = X.DeviceArray.from_host(b), X.DeviceArray(N ** 3 * nv)
    S.precond_apply(db, dx)
    assert np.array_equal(dx.to_host(), O.mg_apply(O.Hier(levels), O.params_from(prm), nv, b))


@pytest.mark.parametrize("mg_prec", [0, 2])
@pytest.mark.parametrize("k", [2, 4])
def test_fused_zero_x_chebyshev_step_is_identical(gpu, k, mg_prec):
    # the pre-smoother's zero-x first step fused into the next SpMV (OP_CHEBZ) against the
    # two launches, every nv; the launches of a solve show which ran
    X = gpu
    N = 12
    levels = levels_for(X, N)
    prm = X.default_params(method="MergedPBiCGStab", smoother="Chebyshev", poly_order=k,
                           sweeps=1, mg_prec=mg_prec)
    try:
        for nv in (1, 2, 4, 8, 16, 32, 64):
            print(f"case k {k} mg_prec {mg_prec} nv {nv}", flush=True)
            b = X.rhs_uniform(nv, N ** 3, nv)
            out, names = [], []
            for fuse in (0, 1):
                X.set_option("cheb_fuse_zero", fuse)
                S = X.Solver(X.Hier(levels), prm, nv)
                db, dx = X.DeviceArray.from_host(b), X.DeviceArray(N ** 3 * nv)
                S.precond_apply(db, dx)
                out.append(dx.to_host())
                dx.zero()
                names.append({q["kernel"] for q in S.profile(db, dx, 1)})
            assert np.array_equal(out[0], out[1]), nv
            assert not any("OP_CHEBZ" in q for q in names[0])
            assert any("OP_CHEBZ" in q for q in names[1]), names[1]
    finally:
        X.set_option("cheb_fuse_zero", 1)


@pytest.mark.parametrize("mg_prec", [0, 2])
def test_hsgs_offblock_matrix_is_identical(gpu, mg_prec):
    # the off-block residual of hybrid GS over the off-block entries kept as a matrix of
    # their own (OP_RESID, the grid path for the 27-point operator) against OP_OFFBLOCK
    # over A, which skips the in-block columns: the same products in the same order
    X = gpu
    ops = (("cube", X.cube(12)[:3]), ("q1fem27", X.q1fem27(12, 12, 12)[:3]))
    try:
        for name, (row, col, val) in ops:
            levels = X.Amg(row, col, val).export()
            n = len(row) - 1
            prm = X.default_params(method="MergedPBiCGStab", smoother="HSGS", sweeps=2,
                                   mg_prec=mg_prec)
            for nv in (1, 4, 8, 32):
                print(f"case {name} mg_prec {mg_prec} nv {nv}", flush=True)
                b = X.rhs_uniform(nv + 3, n, nv)
                out, names = [], []
                for on in (0, 1):
                    X.set_option("hsgs_offblock_csr", on)
                    S = X.Solver(X.Hier(levels), prm, nv)
                    db, dx = X.DeviceArray.from_host(b), X.DeviceArray(n * nv)
                    S.precond_apply(db, dx)
                    out.append(dx.to_host())
                    dx.zero()
                    names.append({q["kernel"] for q in S.profile(db, dx, 1)})
                assert np.array_equal(out[0], out[1]), (name, nv)
                assert any("OP_OFFBLOCK" in q for q in names[0]), names[0]
                assert not any("OP_OFFBLOCK" in q for q in names[1]), names[1]
    finally:
        X.set_option("hsgs_offblock_csr", 1)


@pytest.mark.parametrize("mg_prec", [0, 2])
def test_hsgs_core_path_bit_exact(gpu, O, mg_prec):
    # blocks small enough for the staged core-block sweep in both precisions (32-row
    # blocks: split_range gives 40 rows); the launches of a solve show that kernel ran
    X = gpu
    N, nv = 20, 8
    levels = levels_for(X, N)
    prm = X.default_params(method="MergedPBiCGStab", smoother="HSGS", sweeps=1, mg_prec=mg_prec)
    prm.hsgs_block_rows = 32
    S = X.Solver(X.Hier(levels), prm, nv)
    b = X.rhs_uniform(9, N ** 3, nv)
    db, dx = X.DeviceArray.from_host(b), X.DeviceArray(N ** 3 * nv)
    S.precond_apply(db, dx)
    assert np.array_equal(dx.to_host(), O.mg_apply(O.Hier(levels), O.params_from(prm), nv, b))
    dx.zero()
    names = {k["kernel"] for k in S.profile(db, dx, 1)}
    assert f"k_hsgs_core<{'double' if mg_prec == 0 else 'float'},{nv}>" in names, names


def test_graph_and_eager_are_identical(gpu):
    X = gpu
    N, nv = 24, 8
    levels = levels_for(X, N)
    b = X.rhs_uniform(6, N ** 3, nv)
    xs = []
    for graph in (0, 1):
        prm = X.default_params(method="MergedPBiCGStab", smoother="Chebyshev", poly_order=4,
                               graph=graph)
        S = X.Solver(X.Hier(levels), prm, nv)
        xs.append((S.solve(b), S.stats()[0]))
        # a second solve on the same object (graph replay) gives the same result
        assert np.array_equal(S.solve(b), xs[-1][0])
    assert xs[0][1] == xs[1][1] and np.array_equal(xs[0][0], xs[1][0])


def test_hsgs_block_drift_against_exact_gauss_seidel(gpu, O):
    # the GPU HSGS uses many core blocks (Jacobi coupling between them); against the
    # reference at one core (exact symmetric GS) the iteration count drifts by at most 1
    X = gpu
    N, nv = 32, 4
    levels = levels_for(X, N)
    b = X.rhs_uniform(0, N ** 3, nv)
    prm = X.default_params(method="PBiCGStab", smoother="HSGS", sweeps=1)
    g = X.Solver(X.Hier(levels), prm, nv)
    g.solve(b)
    it_gpu = g.stats()[0]
    p1 = X.default_params(method="PBiCGStab", smoother="HSGS", sweeps=1, hsgs_block_rows=0)
    r1 = O.solve(O.Hier(levels), O.params_from(p1), nv, b)
    assert abs(it_gpu - r1["iters"]) <= 1


def test_uneven_convergence_and_absolute_tolerance(gpu, O):
    X = gpu
    N, nv = 16, 4
    levels = levels_for(X, N)
    b = (X.rhs_uniform(1, N ** 3, nv).reshape(-1, nv) * (10.0 ** -np.arange(nv) * 3)).ravel()
    prm = X.default_params(method="PBiCGStab", smoother="Jacobi", sweeps=2, rel_tol=0.0,
                           abs_tol=1e-6, max_iters=60)
    prm.pre.relax_factor = prm.post.relax_factor = 0.7
    g, r = run_both(X, O, levels, prm, nv, b, hist=True)
    assert abs(g["iters"] - r["iters"]) <= 1 and g["conv"].all()
    assert rel_diff(g["x"], r["x"], nv) <= 1e-6
    np.testing.assert_allclose(g["abs_res"], r["abs_res"], rtol=1e-6)


def test_nonzero_initial_guess_and_max_iters(gpu, O):
    X = gpu
    N, nv = 16, 2
    levels = levels_for(X, N)
    b = X.rhs_uniform(3, N ** 3, nv)
    x0 = X.rhs_uniform(4, N ** 3, nv)
    prm = X.default_params(method="PBiCGStab", smoother="Chebyshev", poly_order=2,
                           max_iters=2, rel_tol=1e-14)
    g, r = run_both(X, O, levels, prm, nv, b, x0=x0)
    assert g["iters"] == r["iters"] == 2 and not g["conv"].any()
    assert rel_diff(g["x"], r["x"], nv) <= 1e-9


def test_breakdown_behaviour_matches_reference(gpu, O):
    # A = 2I: omega = 0/0 after the first half-step; the reference reports every RHS
    # converged with a NaN solution (no breakdown guard) -- the GPU does the same
    X = gpu
    n, nv = 64, 2
    row = np.arange(n + 1, dtype=np.uint32)
    col = np.arange(n, dtype=np.uint32)
    levels = [{"A": (n, n, row, col, np.full(n, 2.0)), "eig": (0.0, 1.0)}]
    prm = X.default_params(method="PBiCGStab", precond="Chebyshev", poly_order=1)
    prm.precond = X.role("Identity")
    b = X.rhs_uniform(5, n, nv)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["iters"] == r["iters"] == 1 and g["conv"].all() and r["conv"].all()
    assert np.isnan(g["x"]).all() and np.isnan(r["x"]).all()


@pytest.mark.parametrize("name", ["Ch", "Ch-rp", "GS", "Jac"])
def test_reference_yaml_configs(gpu, O, name):
    # the reference's single-node parameter sets: PBiCGStab, exactly 20 iterations
    # (tolerances 0), MG with Chebyshev / reduced precision / HSGS / Jacobi, Direct coarse
    X = gpu
    roles = yaml_params.load(os.path.join(GOLD, "yaml", f"config-PBiCGStab-MG-{name}.yaml"))
    N, nv = 24, 4
    row, col, val = X.cube(N)
    levels = X.Amg(row, col, val, **yaml_params.amg_kwargs(roles)).export()
    prm = yaml_params.to_params(X, roles)
    b = X.rhs_uniform(8, N ** 3, nv)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["iters"] == r["iters"] == 20
    xr = r["x"].reshape(-1, nv)
    assert np.isfinite(g["x"]).all() == np.isfinite(r["x"]).all()
    if np.isfinite(xr).all():
        assert rel_diff(g["x"], r["x"], nv) <= 1e-6


@pytest.mark.parametrize("method", ["PBiCGStab", "MergedPBiCGStab", "BiCGStab"])
@pytest.mark.parametrize("graph", [0, 1])
def test_true_residual_replacement_parity(gpu, O, method, graph):
    # every 2nd iteration r = b - A x (fp64) and rho = (r, r0), on the GPU (a second
    # captured graph on those iterations) and in the oracle
    # (the unpreconditioned case at the size and tolerance of the unpreconditioned parity
    # test: over many unpreconditioned iterations the dot reduction order moves the count)
    X = gpu
    pre = "P" in method
    N, nv = (20, 4) if pre else (12, 4)
    levels = levels_for(X, N)
    b = X.rhs_uniform(21, N ** 3, nv)
    prm = X.default_params(method=method, smoother="Chebyshev", poly_order=3, graph=graph,
                           true_res_period=2, rel_tol=1e-10 if pre else 1e-8, max_iters=400,
                           precond="MultiGrid" if pre else None)
    out, ref = run_both(X, O, levels, prm, nv, b)
    assert out["conv"].all() and ref["conv"].all()
    assert abs(out["iters"] - ref["iters"]) <= 1
    assert rel_diff(out["x"], ref["x"], nv) <= 1e-6


def test_true_residual_replacement_rejected_for_merged_bicgstab(gpu):
    X = gpu
    levels = levels_for(X, 8)
    prm = X.default_params(method="MergedBiCGStab", precond=None, true_res_period=1)
    with pytest.raises(X.XamgError):
        X.Solver(X.Hier(levels), prm, 1)


@pytest.mark.parametrize("method", ["PCG", "RBiCGStab", "MergedRBiCGStab"])
def test_true_residual_replacement_rejected_for_other_methods(gpu, method):
    X = gpu
    levels = levels_for(X, 8)
    prm = X.default_params(method=method, true_res_period=2)
    with pytest.raises(X.XamgError):
        X.Solver(X.Hier(levels), prm, 1)


@pytest.mark.parametrize("method", ["PCG", "RBiCGStab", "MergedRBiCGStab"])
def test_new_methods_uneven_convergence_and_max_iters(gpu, O, method):
    # absolute tolerance: RHS of different scale converge at different iterations (iconv
    # masking of the x update); then a max_iters stop before convergence
    X = gpu
    N, nv = 16, 4
    levels = levels_for(X, N)
    b = X.rhs_uniform(3, N ** 3, nv).reshape(-1, nv) * np.array([1.0, 1e-2, 1e2, 1e-4])
    b = b.ravel()
    prm = X.default_params(method=method, smoother="Jacobi", sweeps=2, rel_tol=0.0,
                           abs_tol=1e-6, max_iters=80)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["conv"].all() and r["conv"].all() and abs(g["iters"] - r["iters"]) <= 1
    assert rel_diff(g["x"], r["x"], nv) <= 1e-6
    prm = X.default_params(method=method, smoother="Jacobi", sweeps=2, rel_tol=1e-14,
                           max_iters=3)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["iters"] == r["iters"] == 3 and not g["conv"].any()
    assert rel_diff(g["x"], r["x"], nv) <= 1e-9
