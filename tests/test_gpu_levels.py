"""Per-level parameter roles on the GPU: smoothers that differ by level (params.get(
"pre_smoother", l), src/solvers/detail/mg.inl:113-131) and per-level AMG setup keys (the
"@levX-Y" overrides of the reference example, examples/c/example.c:182-190), checked
against the oracle on the same hierarchy, through the xh_* ABI and the XAMG_* C API."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_capi import api as capi
from test_gpu_capi import new
from test_gpu_solver import levels_for, rel_diff, run_both

pytestmark = pytest.mark.gpu


def level_params(X, nl, **kw):
    """Chebyshev-4 on level 0, Jacobi (2 sweeps) on levels >= 1, HSGS post-smoothing on
    level 1"""
    prm = X.default_params(method=kw.pop("method", "MergedPBiCGStab"), smoother="Chebyshev",
                           poly_order=4, **kw)
    jac = X.role("Jacobi", max_iters=2, relax_factor=0.8)
    hs = X.role("HSGS", max_iters=1)
    pre = [None] + [jac] * (nl - 1)
    post = [None, hs] + [jac] * (nl - 2)
    return X.set_level_roles(prm, pre, post)


@pytest.mark.parametrize("mg_prec", [0, 1, 2])
def test_per_level_smoothers_bit_exact_preconditioner(gpu, O, mg_prec):
    X = gpu
    N, nv = 20, 8
    levels = levels_for(X, N)
    prm = level_params(X, len(levels), mg_prec=mg_prec)
    S = X.Solver(X.Hier(levels), prm, nv)
    b = X.rhs_uniform(9, N ** 3, nv)
    db, dx = X.DeviceArray.from_host(b), X.DeviceArray(N ** 3 * nv)
    S.precond_apply(db, dx)
    ref = O.mg_apply(O.Hier(levels), O.params_from(prm), nv, b)
    assert np.array_equal(dx.to_host(), ref)
    # and it is not the list-wide configuration
    p0 = X.default_params(method="MergedPBiCGStab", smoother="Chebyshev", poly_order=4,
                          mg_prec=mg_prec)
    assert not np.array_equal(ref, O.mg_apply(O.Hier(levels), O.params_from(p0), nv, b))


@pytest.mark.parametrize("method", ["MergedPBiCGStab", "PCG"])
def test_per_level_smoothers_solve_parity(gpu, O, method):
    X = gpu
    N, nv = 24, 4
    levels = levels_for(X, N)
    prm = level_params(X, len(levels), method=method)
    b = X.rhs_uniform(3, N ** 3, nv)
    g, r = run_both(X, O, levels, prm, nv, b)
    assert g["conv"].all() and r["conv"].all()
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel_diff(g["x"], r["x"], nv) <= 1e-6


def test_unsupported_level_role_is_rejected(gpu):
    X = gpu
    levels = levels_for(X, 12)
    prm = X.default_params(method="PBiCGStab", smoother="Chebyshev", poly_order=2)
    X.set_level_roles(prm, [None, X.role("Chebyshev", poly_order=40)])
    with pytest.raises(X.XamgError):
        X.Solver(X.Hier(levels), prm, 2)


def test_c_api_overrides_of_the_reference_example(gpu, O):
    # the call sequence of examples/c/example.c with its overrides: per-level hierarchy,
    # PMIS coarsening and extended+i interpolation on levels 0-1, plus Jacobi smoothing
    # from level 1 on; the solve matches the oracle on the hierarchy the same keys give
    X = gpu
    h = capi(X)
    h.XAMG_param_override_create.argtypes = [C.POINTER(C.c_size_t)]
    h.XAMG_param_override_add_value.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t]
    h.XAMG_param_override_apply.argtypes = [C.c_char_p, C.c_size_t, C.c_size_t]
    h.XAMG_param_override_destroy.argtypes = [C.c_size_t]
    assert h.XAMG_init(0, None, b"") == 0
    N, nv = 20, 2
    row, col, val = X.cube(N)
    n = N ** 3
    cid = new(h, "XAMG_csr_matrix_create")
    assert h.XAMG_csr_matrix_alloc(n, len(col), cid) == 0
    r32, c32 = row.astype(np.int32), col.astype(np.int32)
    assert h.XAMG_csr_matrix_fill(r32.ctypes.data, c32.ctypes.data, val.ctypes.data, cid) == 0
    mid = new(h, "XAMG_matrix_create")
    assert h.XAMG_matrix_construct(cid, mid) == 0
    pid = new(h, "XAMG_param_create")
    for role, key, v in ((b"solver", b"method", b"PBiCGStab"),
                         (b"preconditioner", b"method", b"MultiGrid"),
                         (b"pre_smoother", b"method", b"Chebyshev"),
                         (b"post_smoother", b"method", b"Chebyshev")):
        assert h.XAMG_param_add_value_s(role, key, v, pid) == 0
    assert h.XAMG_param_add_value_i(b"preconditioner", b"hypre_per_level_hierarchy", 1,
                                    pid) == 0
    assert h.XAMG_param_add_value_i(b"preconditioner", b"mg_agg_num_levels", 0, pid) == 0
    assert h.XAMG_param_add_value_d(b"solver", b"abs_tolerance", 0.0, pid) == 0
    for role, pairs in ((b"preconditioner", ((b"mg_coarsening_type", b"8@lev0-1"),
                                             (b"mg_interpolation_type", b"6@lev0-1"))),
                        (b"pre_smoother", ((b"method", b"Jacobi@lev1-E"),
                                           (b"max_iters", b"2@lev1-E"))),
                        (b"post_smoother", ((b"method", b"Jacobi@lev1-E"),
                                            (b"max_iters", b"2@lev1-E")))):
        ov = new(h, "XAMG_param_override_create")
        for k, v in pairs:
            assert h.XAMG_param_override_add_value(k, v, ov) == 0
        assert h.XAMG_param_override_apply(role, pid, ov) == 0
        h.XAMG_param_override_destroy(ov)
    h.XAMG_param_set_default(pid)
    xid = new(h, "XAMG_vector_create", 0)
    bid = new(h, "XAMG_vector_create", 0)
    assert h.XAMG_vector_alloc_d(n, nv, xid) == 0 and h.XAMG_vector_alloc_d(n, nv, bid) == 0
    b = X.rhs_uniform(12, n, nv)
    assert h.XAMG_vector_upload_d(b.ctypes.data, n, 0, nv, bid) == 0
    sid = C.c_size_t()
    assert h.XAMG_solver_create2(mid, pid, xid, bid, C.byref(sid)) == 0
    assert h.XAMG_solver_solve2(nv, xid, bid, sid) == 0
    x = np.zeros(n * nv)
    assert h.XAMG_vector_download_d(x.ctypes.data, n, 0, nv, xid) == 0
    it = C.c_int()
    assert h.XAMG_solver_get_stats(sid, C.byref(it), None, None) == 0
    # oracle on the hierarchy of the same keys
    lk = [dict(coarsen_type=8, interp_type=6)] * 2
    levels = X.Amg(row, col, val, agg_num_levels=0, level_keys=lk).export()
    prm = X.default_params(method="PBiCGStab", smoother="Chebyshev", poly_order=2,
                           abs_tol=0.0)
    jac = X.role("Jacobi", max_iters=2)
    X.set_level_roles(prm, [None] + [jac] * (X.MAX_LEVELS - 1),
                      [None] + [jac] * (X.MAX_LEVELS - 1))
    ref = O.solve(O.Hier(levels), O.params_from(prm), nv, b)
    assert ref["conv"].all() and abs(it.value - ref["iters"]) <= 1
    assert rel_diff(x, ref["x"], nv) <= 1e-6
    assert h.XAMG_solver_destroy(sid) == 0
