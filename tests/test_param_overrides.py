"""Level overrides of the XAMG_* parameter API ("key=value@levX[-Y]", api/c/xamg_c_api.cpp:
256-260, examples/common/cmdline/cmdline.h:79-97) resolved into per-level smoother roles
(params.get("pre_smoother", l), src/solvers/detail/mg.inl:113-131) and per-level AMG setup
keys (hypre_per_level_hierarchy, src/hypre/detail/hypre_hierarchy.inl:54-108).  CPU only:
the parameter calls do not touch the GPU."""
import ctypes as C

import pytest

S = C.c_size_t


def api(X):
    h = X.host()
    h.XAMG_param_create.argtypes = [C.POINTER(S)]
    h.XAMG_param_add_value_s.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, S]
    h.XAMG_param_add_value_i.argtypes = [C.c_char_p, C.c_char_p, C.c_uint16, S]
    h.XAMG_param_set_default.argtypes = [S]
    h.XAMG_param_set_default.restype = None
    h.XAMG_param_override_create.argtypes = [C.POINTER(S)]
    h.XAMG_param_override_add_value.argtypes = [C.c_char_p, C.c_char_p, S]
    h.XAMG_param_override_apply.argtypes = [C.c_char_p, S, S]
    h.XAMG_param_override_destroy.argtypes = [S]
    h.xamg_param_resolve.argtypes = [S, C.c_void_p, C.c_void_p, C.c_void_p]
    h.xamg_param_resolve.restype = C.c_int
    return h


def params(h, entries):
    pid = S()
    assert h.XAMG_param_create(C.byref(pid)) == 0
    for role, key, v in entries:
        if isinstance(v, int):
            assert h.XAMG_param_add_value_i(role.encode(), key.encode(), v, pid) == 0
        else:
            assert h.XAMG_param_add_value_s(role.encode(), key.encode(), v.encode(), pid) == 0
    return pid.value


def override(h, pid, role, pairs):
    ov = S()
    assert h.XAMG_param_override_create(C.byref(ov)) == 0
    for k, v in pairs:
        assert h.XAMG_param_override_add_value(k.encode(), v.encode(), ov) == 0
    rc = h.XAMG_param_override_apply(role.encode(), pid, ov)
    h.XAMG_param_override_destroy(ov)
    return rc


def resolve(X, h, pid):
    p = X.Params()
    a = X.AmgParams()
    lev = (X.AmgLevelKeys * X.MAX_LEVELS)()
    rc = h.xamg_param_resolve(pid, C.byref(p), C.byref(a), C.cast(lev, C.c_void_p))
    return rc, p, a, lev


BASE = [("solver", "method", "MergedPBiCGStab"), ("preconditioner", "method", "MultiGrid"),
        ("pre_smoother", "method", "Chebyshev"), ("post_smoother", "method", "Chebyshev"),
        ("pre_smoother", "polynomial_order", 4), ("post_smoother", "polynomial_order", 4)]


def test_smoother_overrides_give_per_level_roles(X):
    h = api(X)
    pid = params(h, BASE)
    # Jacobi with 2 sweeps from level 1 to the last; order 2 on level 3 only
    assert override(h, pid, "pre_smoother", [("method", "Jacobi@lev1-E"),
                                             ("max_iters", "2@lev1-E")]) == 0
    assert override(h, pid, "post_smoother", [("polynomial_order", "2@lev3")]) == 0
    h.XAMG_param_set_default(pid)
    rc, p, _, _ = resolve(X, h, pid)
    assert rc == 0
    assert p.nlev_roles == X.MAX_LEVELS  # "E": every level
    J, CH = X.METHODS["Jacobi"], X.METHODS["Chebyshev"]
    assert p.pre.method == CH and p.pre.poly_order == 4
    assert p.pre_lev[0].method == CH and p.pre_lev[0].poly_order == 4
    for l in range(1, X.MAX_LEVELS):
        assert p.pre_lev[l].method == J and p.pre_lev[l].max_iters == 2
        # a method switched by an override gets its family's defaults (relax 0.99)
        assert abs(p.pre_lev[l].relax_factor - 0.99) < 1e-7
    assert [p.post_lev[l].poly_order for l in range(5)] == [4, 4, 4, 2, 4]
    assert all(p.post_lev[l].method == CH for l in range(X.MAX_LEVELS))


def test_single_level_and_ranges(X):
    h = api(X)
    pid = params(h, BASE)
    assert override(h, pid, "pre_smoother", [("polynomial_order", "3@lev2-4")]) == 0
    h.XAMG_param_set_default(pid)
    rc, p, _, _ = resolve(X, h, pid)
    assert rc == 0 and p.nlev_roles == 5
    assert [p.pre_lev[l].poly_order for l in range(5)] == [4, 4, 3, 3, 3]
    assert p.pre.poly_order == 4  # levels >= 5 use the list-wide role


def test_setup_overrides_of_the_reference_example(X):
    # examples/c/example.c:182-190: per-level hierarchy, PMIS and extended+i on levels 0-1
    h = api(X)
    pid = params(h, BASE + [("preconditioner", "hypre_per_level_hierarchy", 1)])
    assert override(h, pid, "preconditioner", [("mg_coarsening_type", "8@lev0-1"),
                                               ("mg_interpolation_type", "6@lev0-1")]) == 0
    h.XAMG_param_set_default(pid)
    rc, _, a, lev = resolve(X, h, pid)
    assert rc == 0
    assert a.coarsen_type == 6 and a.interp_type == 0 and a.nlev_keys == 2
    assert [(lev[l].coarsen_type, lev[l].interp_type) for l in range(2)] == [(8, 6), (8, 6)]


def test_setup_overrides_need_a_per_level_hierarchy(X):
    # with the full hierarchy the reference reads the list without a level
    # (hypre_hierarchy.inl:31-52): the overrides do not reach the setup
    h = api(X)
    pid = params(h, BASE)
    assert override(h, pid, "preconditioner", [("mg_coarsening_type", "8@lev0-1")]) == 0
    h.XAMG_param_set_default(pid)
    rc, _, a, _ = resolve(X, h, pid)
    assert rc == 0 and a.nlev_keys == 0 and a.coarsen_type == 6


@pytest.mark.parametrize("role,key,val,code", [
    ("pre_smoother", "method", "Jacobi", -1),            # no level range
    ("pre_smoother", "method", "Jacobi@levX", -1),       # malformed range
    ("pre_smoother", "method", "Jacobi@lev3-1", -1),     # empty range
    ("pre_smoother", "mg_cycle", "1@lev0", -3),          # not a smoother key
    ("coarse_grid_solver", "method", "Jacobi@lev0", -3), # no per-level coarse role
    ("solver", "max_iters", "5@lev0", -3),
    ("preconditioner", "mg_max_levels", "3@lev0", -3),   # not a per-level setup key
])
def test_overrides_that_cannot_be_honoured_are_rejected(X, role, key, val, code):
    h = api(X)
    pid = params(h, BASE)
    assert override(h, pid, role, [(key, val)]) == code
    # a rejected set stores nothing
    h.XAMG_param_set_default(pid)
    rc, p, _, _ = resolve(X, h, pid)
    assert rc == 0 and p.nlev_roles == 0


@pytest.mark.parametrize("key,val", [("mg_coarsening_type", "10"),
                                     ("mg_interpolation_type", "3"),
                                     ("mg_agg_interpolation_type", "1")])
def test_unsupported_setup_types_fail_at_resolve(X, key, val):
    h = api(X)
    pid = params(h, BASE + [("preconditioner", key, int(val))])
    h.XAMG_param_set_default(pid)
    rc, _, _, _ = resolve(X, h, pid)
    assert rc == -3


def test_smoother_convergence_check_is_rejected(X):
    h = api(X)
    pid = params(h, BASE + [("pre_smoother", "convergence_check", 1)])
    h.XAMG_param_set_default(pid)
    assert resolve(X, h, pid)[0] == -3
