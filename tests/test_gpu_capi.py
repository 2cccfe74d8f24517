"""The XAMG_* C API (drop-in for api/c/xamg_c_api.h) end to end on the GPU, following the
call sequence of the reference's examples/c/example.c:142-253, checked against the
oracle on the same hierarchy and against a direct solve."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu

S = C.c_size_t


def api(X):
    h = X.host()
    h.XAMG_init.argtypes = [C.c_int, C.c_void_p, C.c_char_p]
    h.XAMG_get_nv.restype = C.c_int
    for f in ("XAMG_param_create", "XAMG_vector_create", "XAMG_csr_matrix_create",
              "XAMG_matrix_create", "XAMG_part_create"):
        getattr(h, f).restype = C.c_int
    h.XAMG_param_add_value_s.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, S]
    h.XAMG_param_add_value_i.argtypes = [C.c_char_p, C.c_char_p, C.c_uint16, S]
    h.XAMG_param_add_value_d.argtypes = [C.c_char_p, C.c_char_p, C.c_float, S]
    h.XAMG_param_change_value_d.argtypes = [C.c_char_p, C.c_char_p, C.c_float, S]
    h.XAMG_param_set_default.argtypes = [S]
    h.XAMG_param_set_default.restype = None
    h.XAMG_csr_matrix_alloc.argtypes = [C.c_uint64, C.c_uint64, S]
    h.XAMG_csr_matrix_fill.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, S]
    h.XAMG_csr_matrix_get_nrows.argtypes = [S]
    h.XAMG_csr_matrix_get_nrows.restype = C.c_uint32
    h.XAMG_matrix_construct.argtypes = [S, S]
    h.XAMG_part_construct.argtypes = [C.c_int, S]
    h.XAMG_part_get_numa_size.argtypes = [S]
    h.XAMG_vector_create.argtypes = [C.POINTER(S), C.c_int]
    h.XAMG_vector_alloc_d.argtypes = [C.c_int, C.c_uint16, S]
    h.XAMG_vector_upload_d.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint16, S]
    h.XAMG_vector_download_d.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint16, S]
    h.XAMG_vector_set_val_d.argtypes = [C.c_double, S]
    h.XAMG_solver_create2.argtypes = [S, S, S, S, C.POINTER(S)]
    h.XAMG_solver_solve1.argtypes = [C.c_uint16, S]
    h.XAMG_solver_solve2.argtypes = [C.c_uint16, S, S, S]
    h.XAMG_solver_get_convergence_info.argtypes = [C.c_void_p, C.c_uint16, S]
    h.XAMG_solver_get_stats.argtypes = [S, C.POINTER(C.c_int), C.c_void_p, C.c_void_p]
    h.XAMG_solver_renew_params.argtypes = [S, S]
    h.XAMG_solver_destroy.argtypes = [S]
    return h


def new(h, f, *extra):
    i = S()
    assert getattr(h, f)(C.byref(i), *extra) == 0
    return i.value


def test_c_api_example_flow(gpu, O):
    X = gpu
    h = api(X)
    assert h.XAMG_init(0, None, b"") == 0
    N, nv = 20, 2
    row, col, val = X.cube(N)
    n = N ** 3
    cid = new(h, "XAMG_csr_matrix_create")
    assert h.XAMG_csr_matrix_alloc(n, len(col), cid) == 0
    r32, c32 = row.astype(np.int32), col.astype(np.int32)
    assert h.XAMG_csr_matrix_fill(r32.ctypes.data, c32.ctypes.data, val.ctypes.data, cid) == 0
    assert h.XAMG_csr_matrix_get_nrows(cid) == n
    mid = new(h, "XAMG_matrix_create")
    pid_part = new(h, "XAMG_part_create")
    assert h.XAMG_part_construct(n, pid_part) == 0 and h.XAMG_part_get_numa_size(pid_part) == n
    assert h.XAMG_matrix_construct(cid, mid) == 0
    pid = new(h, "XAMG_param_create")
    for role, key, v in ((b"solver", b"method", b"PBiCGStab"),
                         (b"preconditioner", b"method", b"MultiGrid"),
                         (b"pre_smoother", b"method", b"Chebyshev"),
                         (b"post_smoother", b"method", b"Chebyshev")):
        assert h.XAMG_param_add_value_s(role, key, v, pid) == 0
    assert h.XAMG_param_add_value_i(b"pre_smoother", b"polynomial_order", 3, pid) == 0
    assert h.XAMG_param_add_value_i(b"post_smoother", b"polynomial_order", 3, pid) == 0
    assert h.XAMG_param_add_value_d(b"solver", b"rel_tolerance", 1e-6, pid) == 0
    assert h.XAMG_param_add_value_d(b"solver", b"abs_tolerance", 0.0, pid) == 0
    h.XAMG_param_set_default(pid)
    xid = new(h, "XAMG_vector_create", 2)
    bid = new(h, "XAMG_vector_create", 2)
    assert h.XAMG_vector_alloc_d(n, nv, xid) == 0 and h.XAMG_vector_alloc_d(n, nv, bid) == 0
    b = X.rhs_uniform(7, n, nv)
    assert h.XAMG_vector_upload_d(b.ctypes.data, n, 0, nv, bid) == 0
    assert h.XAMG_vector_set_val_d(0.0, xid) == 0
    sid = S()
    assert h.XAMG_solver_create2(mid, pid, xid, bid, C.byref(sid)) == 0, \
        h.XAMG_get_last_error()
    sid = sid.value
    assert h.XAMG_solver_solve1(nv, sid) == 0
    conv = np.zeros(nv, np.int32)
    assert h.XAMG_solver_get_convergence_info(conv.ctypes.data, nv, sid) == 0
    assert conv.all()
    it = C.c_int()
    a, r = np.zeros(nv), np.zeros(nv)
    assert h.XAMG_solver_get_stats(sid, C.byref(it), a.ctypes.data, r.ctypes.data) == 0
    assert (r <= np.float32(1e-6)).all()
    x = np.zeros(n * nv)
    assert h.XAMG_vector_download_d(x.ctypes.data, n, 0, nv, xid) == 0
    A = sp.csr_matrix((val.copy(), col.astype(np.int64), row.astype(np.int64)), shape=(n, n))
    xs = spla.spsolve(A.tocsc(), b.reshape(-1, nv))
    err = np.abs(x.reshape(-1, nv) - xs).max(0) / np.abs(xs).max(0)
    assert err.max() < 1e-4
    it1 = it.value
    # tighten the tolerance and solve again from zero (renew_params, example.c:215-223)
    assert h.XAMG_param_change_value_d(b"solver", b"rel_tolerance", 1e-10, pid) == 0
    assert h.XAMG_solver_renew_params(pid, sid) == 0
    assert h.XAMG_vector_set_val_d(0.0, xid) == 0
    assert h.XAMG_solver_solve2(nv, xid, bid, sid) == 0
    assert h.XAMG_solver_get_stats(sid, C.byref(it), a.ctypes.data, r.ctypes.data) == 0
    assert it.value > it1 and (r <= np.float32(1e-10)).all()
    # a wrong nv is an error code, not an abort
    assert h.XAMG_solver_solve1(nv + 1, sid) < 0
    assert h.XAMG_solver_destroy(sid) == 0


@pytest.mark.parametrize("method", ["PCG", "RBiCGStab", "MergedRBiCGStab"])
def test_c_api_system_file_and_registry_methods(gpu, tmp_path, method):
    # a system file in the reference's reader layout (written here, read back per rank)
    # feeds the C API; the solver role names one of the registry's other Krylov methods
    X = gpu
    h = api(X)
    assert h.XAMG_init(0, None, b"") == 0
    N, nv = 16, 2
    row, col, val = X.cube(N)
    n = N ** 3
    b = X.rhs_uniform(11, n, nv)
    path = str(tmp_path / "cube16.bin")
    X.write_system(path, row, col, val, nv, b)
    sysf = X.read_system(path, nv)
    assert sysf["found"] == 1 and (sysf["x"] == 0).all()
    cid = new(h, "XAMG_csr_matrix_create")
    assert h.XAMG_csr_matrix_alloc(n, len(sysf["col"]), cid) == 0
    r32, c32 = sysf["row"].astype(np.int32), sysf["col"].astype(np.int32)
    v = np.ascontiguousarray(sysf["val"])
    assert h.XAMG_csr_matrix_fill(r32.ctypes.data, c32.ctypes.data, v.ctypes.data, cid) == 0
    mid = new(h, "XAMG_matrix_create")
    assert h.XAMG_matrix_construct(cid, mid) == 0
    pid = new(h, "XAMG_param_create")
    for role, key, val_s in ((b"solver", b"method", method.encode()),
                             (b"preconditioner", b"method", b"MultiGrid"),
                             (b"pre_smoother", b"method", b"Jacobi"),
                             (b"post_smoother", b"method", b"Jacobi")):
        assert h.XAMG_param_add_value_s(role, key, val_s, pid) == 0
    assert h.XAMG_param_add_value_d(b"solver", b"rel_tolerance", 1e-9, pid) == 0
    assert h.XAMG_param_add_value_d(b"solver", b"abs_tolerance", 0.0, pid) == 0
    h.XAMG_param_set_default(pid)
    xid = new(h, "XAMG_vector_create", 2)
    bid = new(h, "XAMG_vector_create", 2)
    assert h.XAMG_vector_alloc_d(n, nv, xid) == 0 and h.XAMG_vector_alloc_d(n, nv, bid) == 0
    bb = np.ascontiguousarray(sysf["b"])
    xx = np.ascontiguousarray(sysf["x"])
    assert h.XAMG_vector_upload_d(bb.ctypes.data, n, 0, nv, bid) == 0
    assert h.XAMG_vector_upload_d(xx.ctypes.data, n, 0, nv, xid) == 0
    sid = S()
    assert h.XAMG_solver_create2(mid, pid, xid, bid, C.byref(sid)) == 0, \
        h.XAMG_get_last_error()
    sid = sid.value
    assert h.XAMG_solver_solve1(nv, sid) == 0
    conv = np.zeros(nv, np.int32)
    assert h.XAMG_solver_get_convergence_info(conv.ctypes.data, nv, sid) == 0
    assert conv.all()
    x = np.zeros(n * nv)
    assert h.XAMG_vector_download_d(x.ctypes.data, n, 0, nv, xid) == 0
    A = sp.csr_matrix((val.copy(), col.astype(np.int64), row.astype(np.int64)), shape=(n, n))
    xs = spla.spsolve(A.tocsc(), b.reshape(-1, nv))
    assert (np.abs(x.reshape(-1, nv) - xs).max(0) / np.abs(xs).max(0)).max() < 1e-7
    assert h.XAMG_solver_destroy(sid) == 0
