"""ctypes bindings of libxamg_hip.so (include/xamg_hip.h) and libxamg.so
(include/xamg_amg.h, include/xamg_c_api.h) for the test suite and the benchmark.

Loads the in-tree shared libraries under lib/ and fails loudly when they are missing:
there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

# method ids, include/xamg_hip.h
IDENTITY, JACOBI, HSGS, CHEBYSHEV, DIRECT, MULTIGRID = 0, 1, 2, 3, 4, 5
BICGSTAB, PBICGSTAB, MERGED_BICGSTAB, MERGED_PBICGSTAB = 6, 7, 8, 9
PCG, RBICGSTAB, MERGED_RBICGSTAB = 10, 11, 12
NONE = -1

METHODS = {
    "Identity": IDENTITY, "Jacobi": JACOBI, "HSGS": HSGS, "Chebyshev": CHEBYSHEV,
    "Direct": DIRECT, "MultiGrid": MULTIGRID, "BiCGStab": BICGSTAB, "PBiCGStab": PBICGSTAB,
    "MergedBiCGStab": MERGED_BICGSTAB, "MergedPBiCGStab": MERGED_PBICGSTAB,
    "PCG": PCG, "RBiCGStab": RBICGSTAB, "MergedRBiCGStab": MERGED_RBICGSTAB,
}


MAX_LEVELS = 32  # XH_MAX_LEVELS


class Role(C.Structure):
    _fields_ = [("method", C.c_int32), ("max_iters", C.c_int32), ("abs_tol", C.c_float),
                ("rel_tol", C.c_float), ("conv_check", C.c_int32), ("relax_factor", C.c_float),
                ("poly_order", C.c_int32), ("spectrum_fraction", C.c_float)]


class Params(C.Structure):
    _fields_ = [("solver", Role), ("precond", Role), ("pre", Role), ("post", Role),
                ("coarse", Role), ("mg_cycle", C.c_int32), ("mg_prec", C.c_int32),
                ("hsgs_ncores", C.c_int32), ("hsgs_block_rows", C.c_int32),
                ("coarse_fp64", C.c_int32), ("index_mode", C.c_int32), ("graph", C.c_int32),
                ("true_res_period", C.c_int32), ("reserved", C.c_int32 * 3),
                ("nlev_roles", C.c_int32), ("pre_lev", Role * MAX_LEVELS),
                ("post_lev", Role * MAX_LEVELS)]


class KernelStat(C.Structure):
    _fields_ = [("kernel", C.c_char * 80), ("level", C.c_int32), ("launches", C.c_int32),
                ("bytes", C.c_double), ("ms", C.c_double)]


class AmgParams(C.Structure):
    _fields_ = [("strength", C.c_float), ("max_row_sum", C.c_float), ("coarse_size", C.c_int32),
                ("max_levels", C.c_int32), ("trunc_factor", C.c_float), ("pmax", C.c_int32),
                ("second_pass", C.c_int32), ("agg_num_levels", C.c_int32),
                ("num_paths", C.c_int32), ("agg_trunc_factor", C.c_float),
                ("agg_pmax", C.c_int32), ("eig_iters", C.c_int32), ("eig_seed", C.c_uint64),
                ("coarsen_type", C.c_int32), ("interp_type", C.c_int32),
                ("agg_interp_type", C.c_int32), ("nlev_keys", C.c_int32),
                ("lev_keys", C.c_void_p)]


class AmgLevelKeys(C.Structure):
    _fields_ = [("strength", C.c_float), ("max_row_sum", C.c_float),
                ("trunc_factor", C.c_float), ("pmax", C.c_int32), ("coarsen_type", C.c_int32),
                ("interp_type", C.c_int32), ("agg_interp_type", C.c_int32),
                ("agg_trunc_factor", C.c_float), ("agg_pmax", C.c_int32),
                ("num_paths", C.c_int32)]


def role(method, max_iters=1, abs_tol=0.0, rel_tol=0.0, conv_check=0, relax_factor=0.99,
         poly_order=2, spectrum_fraction=0.3):
    """One parameter role with the reference defaults of a non-solver list."""
    m = METHODS[method] if isinstance(method, str) else method
    return Role(m, max_iters, abs_tol, rel_tol, conv_check, relax_factor, poly_order,
                spectrum_fraction)


_hip = None
_host = None


class XamgError(RuntimeError):
    pass


def _load():
    global _hip, _host
    if _hip is not None:
        return
    hip_path = os.path.join(LIBDIR, "libxamg_hip.so")
    host_path = os.path.join(LIBDIR, "libxamg.so")
    for p in (hip_path, host_path):
        if not os.path.exists(p):
            raise XamgError(f"{p} is not built: run __graft_entry__.build()")
    _hip = C.CDLL(hip_path, mode=C.RTLD_GLOBAL)
    _host = C.CDLL(host_path, mode=C.RTLD_GLOBAL)
    vp, i32, i64, f64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint64
    P = C.POINTER
    _hip.xh_last_error.restype = C.c_char_p
    sig = {
        "xh_init": [C.c_int], "xh_finalize": [], "xh_malloc": [P(vp), C.c_size_t],
        "xh_free": [vp], "xh_memcpy_h2d": [vp, vp, C.c_size_t],
        "xh_fill_zero": [vp, C.c_size_t], "xh_copy_device": [vp, vp, C.c_size_t],
        "xh_memcpy_d2h": [vp, vp, C.c_size_t], "xh_synchronize": [],
        "xh_device_info": [C.c_char_p, C.c_int, P(C.c_int), P(f64)],
        "xh_matrix_create": [i64, i64, i64, vp, vp, vp, C.c_int, C.c_int, P(vp)],
        "xh_matrix_info": [vp, P(i64), P(i64), P(i64), P(C.c_int), P(C.c_int), P(i64),
                           P(C.c_size_t)],
        "xh_matrix_destroy": [vp],
        "xh_spmv": [vp, C.c_int, vp, vp, C.c_int],
        "xh_spmv_timed": [vp, C.c_int, vp, vp, C.c_int, C.c_int, P(f64)],
        "xh_dense_matvec": [i64, C.c_int, C.c_int, vp, vp, vp],
        "xh_matrix_fold": [vp, i64, P(i64)],
        "xh_csr_fold": [i64, i64, vp, vp, P(i64)],
        "xh_csr_grid": [i64, vp, vp, vp, vp, P(i64)],
        "xh_grid_schedule": [C.c_int32, C.c_int32, C.c_int32, C.c_int, C.c_int, C.c_int,
                             P(i64), P(i64), P(i64), vp],
        "xh_matrix_grid": [vp, vp],
        "xh_matrix_layouts": [vp, P(i32)],
        "xh_spmv_schedule": [i64, i64, i64, C.c_int, C.c_int, C.c_int, P(i64), P(i64), P(i64),
                             vp],
        "xh_set_option": [C.c_char_p, i64],
        "xh_matrix_create_compressed": [i64, i64, i64, i64, vp, vp, C.c_int, vp, vp, i64, i64,
                                        vp, C.c_int, P(vp)],
        "xh_matrix_gather": [vp, C.c_int, vp, vp],
        "xh_batch_create": [C.c_int, vp, vp, vp, vp, vp, vp, P(Params), P(vp)],
        "xh_batch_set_rhs": [vp, C.c_int, vp],
        "xh_batch_solve": [vp],
        "xh_batch_get_x": [vp, C.c_int, vp],
        "xh_batch_stats": [vp, C.c_int, P(i32), vp, vp, vp, P(f64)],
        "xh_batch_autotune": [vp, C.c_int, vp],
        "xh_batch_spmv_timed": [vp, C.c_int, P(f64), P(f64)],
        "xh_batch_set_tiles": [vp, vp],
        "xh_batch_set_persistent": [vp, C.c_int],
        "xh_batch_set_layout": [vp, vp, vp],
        "xh_batch_get_layout": [vp, vp, vp],
        "xh_batch_destroy": [vp],
        "xh_unet_create": [C.c_int, vp, u64, P(vp)],
        "xh_unet_destroy": [vp],
        "xh_unet_nparams": [vp, P(i64)],
        "xh_unet_get_params": [vp, vp],
        "xh_unet_set_params": [vp, vp],
        "xh_unet_get_output_scale": [vp, P(C.c_float)],
        "xh_unet_set_output_scale": [vp, C.c_float],
        "xh_unet_loss_grad": [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, P(C.c_float), vp,
                              vp],
        "xh_unet_train": [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int,
                          C.c_int, C.c_int, C.c_float, u64, vp],
        "xh_unet_infer": [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp],
        "xh_unet_mfma_selftest": [vp, vp, vp],
        "xh_eval_bumps": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp],
        "xh_eval_bumps_noalloc": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp],
        "xh_hier_create": [C.c_int, P(vp)],
        "xh_hier_set_level": [vp, C.c_int, i64, i64, vp, vp, vp, i64, i64, vp, vp, vp, i64, vp,
                              vp, vp, f64, f64],
        "xh_hier_destroy": [vp],
        "xh_solver_create": [vp, P(Params), C.c_int, P(vp)],
        "xh_solver_solve": [vp, vp, vp], "xh_solver_solve_device": [vp, vp, vp],
        "xh_solver_stats": [vp, P(i32), vp, vp, vp],
        "xh_solver_history": [vp, vp, i32, P(i32)],
        "xh_solver_precond_apply": [vp, vp, vp],
        "xh_solver_profile": [vp, vp, vp, i32, vp, i32, P(i32)],
        "xh_solver_smooth": [vp, vp, vp, i32],
        "xh_copy_device_timed": [vp, vp, C.c_size_t, C.c_int, P(f64)],
        "xh_solver_renew": [vp, P(Role)],
        "xh_solver_info": [vp, P(C.c_int), P(C.c_size_t)],
        "xh_solver_destroy": [vp],
    }
    for name, args in sig.items():
        f = getattr(_hip, name)
        f.argtypes = args
        f.restype = C.c_int
    hsig = {
        "xamg_amg_setup": [i64, vp, vp, vp, P(AmgParams), P(vp)],
        "xamg_amg_nlevels": [vp],
        "xamg_amg_level": [vp, C.c_int, C.c_int, P(i64), P(i64), P(i64), P(vp), P(vp), P(vp)],
        "xamg_amg_eigs": [vp, C.c_int, P(f64), P(f64)],
        "xamg_amg_estimate_eigs": [vp, C.c_int, u64],
        "xamg_amg_to_device": [vp, P(vp)],
        "xamg_amg_free": [vp],
        "xamg_amg_strength": [i64, vp, vp, vp, C.c_float, C.c_float, vp],
        "xamg_amg_split": [i64, vp, vp, vp, vp, C.c_int, u64, vp],
        "xamg_amg_interp": [i64, vp, vp, vp, vp, vp, C.c_int, C.c_float, C.c_int, P(vp)],
        "xamg_cube": [i64, i64, i64, vp, vp, vp],
        "xamg_lognormal_field": [u64, f64, i64, vp],
        "xamg_varcoef7": [i64, i64, i64, vp, vp, vp, vp],
        "xamg_q1fem27": [i64, i64, i64, vp, vp, vp, vp],
        "xamg_rhs_bumps": [u64, i64, i64, i64, C.c_int, C.c_int, C.c_int, f64, f64, vp],
        "xamg_bump_params": [u64, i64, i64, i64, C.c_int, C.c_int, C.c_int, f64, f64, vp, vp],
        "xamg_rhs_uniform": [u64, i64, vp],
    }
    for name, args in hsig.items():
        f = getattr(_host, name)
        f.argtypes = args
        f.restype = C.c_int
    for f in ("xamg_cube_nnz", "xamg_varcoef7_nnz", "xamg_q1fem27_nnz"):
        getattr(_host, f).argtypes = [i64, i64, i64]
        getattr(_host, f).restype = i64
    cp = C.c_char_p
    for name, args in {
        "xamg_io_read_header": [cp, vp, vp],
        "xamg_io_sections": [cp, C.c_int, vp, vp],
        "xamg_io_block_nnz": [cp, i64, i64, vp],
        "xamg_io_read_system": [cp, C.c_int, i64, i64, vp, vp, vp, vp, vp, vp],
        "xamg_io_write_system": [cp, i64, i64, vp, vp, vp, C.c_int, vp, vp],
    }.items():
        f = getattr(_host, name)
        f.argtypes = args
        f.restype = C.c_int
    _host.xamg_io_block_range.argtypes = [i64, C.c_int, C.c_int, vp, vp]
    _host.xamg_io_block_range.restype = None
    _host.xamg_amg_default_params.argtypes = [P(AmgParams)]
    _host.xamg_amg_default_params.restype = None
    _host.XAMG_get_last_error.restype = C.c_char_p


def hip():
    _load()
    return _hip


def host():
    _load()
    return _host


def check(rc, what=""):
    if rc != 0:
        msg = _hip.xh_last_error().decode() if _hip is not None else ""
        hmsg = _host.XAMG_get_last_error().decode() if _host is not None else ""
        raise XamgError(f"{what} failed ({rc}): {msg or hmsg}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------- host-side helpers ----------------

def cube(nx, ny=None, nz=None):
    """7-point cube Laplacian of the reference generator as (row, col, val) numpy arrays."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    h = host()
    nnz = h.xamg_cube_nnz(nx, ny, nz)
    n = nx * ny * nz
    row = np.zeros(n + 1, np.uint32)
    col = np.zeros(nnz, np.uint32)
    val = np.zeros(nnz, np.float64)
    check(h.xamg_cube(nx, ny, nz, _ptr(row), _ptr(col), _ptr(val)), "xamg_cube")
    return row, col, val


def lognormal_field(seed, sigma, count):
    k = np.zeros(count)
    check(host().xamg_lognormal_field(seed, sigma, count, _ptr(k)), "xamg_lognormal_field")
    return k


def varcoef7(nx, ny=None, nz=None, sigma=1.0, seed=0, kcell=None):
    """-div(k grad u), 7-point finite volumes, lognormal cell k: (row, col, val, k)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = nx * ny * nz
    k = lognormal_field(seed, sigma, n) if kcell is None else np.ascontiguousarray(kcell, float)
    nnz = host().xamg_varcoef7_nnz(nx, ny, nz)
    row, col, val = np.zeros(n + 1, np.uint32), np.zeros(nnz, np.uint32), np.zeros(nnz)
    check(host().xamg_varcoef7(nx, ny, nz, _ptr(k), _ptr(row), _ptr(col), _ptr(val)),
          "xamg_varcoef7")
    return row, col, val, k


def io_block_range(nrows, nprocs, proc):
    """the reference's per-rank row block (include/xamg_io.h): (row_begin, block_rows)"""
    rb, nr = C.c_int64(), C.c_int64()
    host().xamg_io_block_range(nrows, nprocs, proc, C.byref(rb), C.byref(nr))
    return rb.value, nr.value


def read_system(path, nv=1, nprocs=1, proc=0):
    """one row block of a binary system file (include/xamg_io.h): dict(row, col, val, b, x,
    nrows, found, row_begin); found bit 0: b from the file, bit 1: x from the file"""
    bp = os.fsencode(path)
    n, nnz = C.c_int64(), C.c_int64()
    check(host().xamg_io_read_header(bp, C.byref(n), C.byref(nnz)), "xamg_io_read_header")
    rb, nr = io_block_range(n.value, nprocs, proc)
    bnnz = C.c_int64()
    check(host().xamg_io_block_nnz(bp, rb, nr, C.byref(bnnz)), "xamg_io_block_nnz")
    row = np.zeros(nr + 1, np.uint32)
    col = np.zeros(max(bnnz.value, 1), np.uint32)
    val = np.zeros(max(bnnz.value, 1))
    b = np.zeros(max(nr * nv, 1))
    x = np.zeros(max(nr * nv, 1))
    found = C.c_int32()
    check(host().xamg_io_read_system(bp, nv, rb, nr, _ptr(row), _ptr(col), _ptr(val), _ptr(b),
                                     _ptr(x), C.byref(found)), "xamg_io_read_system")
    k = bnnz.value
    return dict(row=row, col=col[:k], val=val[:k], b=b[:nr * nv], x=x[:nr * nv],
                nrows=n.value, found=found.value, row_begin=rb)


def write_system(path, row, col, val, nv=1, b=None, x=None):
    row = np.ascontiguousarray(row, np.uint32)
    col = np.ascontiguousarray(col, np.uint32)
    val = np.ascontiguousarray(val, np.float64)
    bb = None if b is None else np.ascontiguousarray(b, np.float64)
    xx = None if x is None else np.ascontiguousarray(x, np.float64)
    check(host().xamg_io_write_system(os.fsencode(path), len(row) - 1, len(col), _ptr(row),
                                      _ptr(col), _ptr(val), nv,
                                      None if bb is None else _ptr(bb),
                                      None if xx is None else _ptr(xx)),
          "xamg_io_write_system")


def q1fem27(nx, ny=None, nz=None, sigma=2.0, seed=0, kcell=None):
    """-div(k grad u), trilinear finite elements (27-point), lognormal cell k:
    (row, col, val, k)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = nx * ny * nz
    nc = (nx + 1) * (ny + 1) * (nz + 1)
    k = lognormal_field(seed, sigma, nc) if kcell is None else np.ascontiguousarray(kcell, float)
    nnz = host().xamg_q1fem27_nnz(nx, ny, nz)
    row, col, val = np.zeros(n + 1, np.uint32), np.zeros(nnz, np.uint32), np.zeros(nnz)
    check(host().xamg_q1fem27(nx, ny, nz, _ptr(k), _ptr(row), _ptr(col), _ptr(val)),
          "xamg_q1fem27")
    return row, col, val, k


def rhs_bumps(seed, nx, ny, nz, nv, min_bumps=8, max_bumps=32, min_width=2.0, max_width=12.0):
    b = np.zeros(nx * ny * nz * nv)
    check(host().xamg_rhs_bumps(seed, nx, ny, nz, nv, min_bumps, max_bumps, min_width,
                                max_width, _ptr(b)), "xamg_rhs_bumps")
    return b


def rhs_bumps_device(seed, nx, ny, nz, nv, min_bumps=8, max_bumps=32, min_width=2.0,
                     max_width=12.0, out=None):
    """the right-hand sides of rhs_bumps, evaluated on the GPU into a DeviceArray"""
    count = np.zeros(nv, np.int32)
    P = np.zeros(nv * max_bumps * 5)
    check(host().xamg_bump_params(seed, nx, ny, nz, nv, min_bumps, max_bumps, min_width,
                                  max_width, _ptr(count), _ptr(P)), "xamg_bump_params")
    out = DeviceArray(nx * ny * nz * nv) if out is None else out
    check(hip().xh_eval_bumps(nx, ny, nz, nv, max_bumps, _ptr(count), _ptr(P), out.ptr),
          "xh_eval_bumps")
    return out


def csr_fold(row, col, ncols=None):
    """The fold (dominant long-range column offset) xh_matrix_create would detect."""
    n = len(row) - 1
    f = C.c_int64()
    check(hip().xh_csr_fold(n, n if ncols is None else ncols, _ptr(row), _ptr(col),
                            C.byref(f)), "xh_csr_fold")
    return f.value


def spmv_schedule(nrows, nnz, fold, nv, prec=0, dots=False):
    """(grid, tiles per workgroup, [ntiles, 2] row ranges in processing order) of an SpMV
    launch."""
    g, nt, tpw = C.c_int64(), C.c_int64(), C.c_int64()
    check(hip().xh_spmv_schedule(nrows, nnz, fold, prec, nv, int(dots), C.byref(g),
                                 C.byref(nt), C.byref(tpw), None), "xh_spmv_schedule")
    rows = np.zeros((nt.value, 2), np.int64)
    check(hip().xh_spmv_schedule(nrows, nnz, fold, prec, nv, int(dots), None, None, None,
                                 rows.ctypes.data), "xh_spmv_schedule")
    return g.value, tpw.value, rows


def csr_grid(row, col):
    """(dims (nx, ny, nz) or None, per-entry grid codes, coded count) of the structured-grid
    view xh_matrix_create would attach (when coded >= 90 % of the entries)"""
    n = len(row) - 1
    row = np.ascontiguousarray(row, np.uint32)
    col = np.ascontiguousarray(col, np.uint32)
    dims = np.zeros(3, np.int32)
    code = np.full(max(len(col), 1), 255, np.uint8)
    coded = C.c_int64(0)
    check(hip().xh_csr_grid(n, _ptr(row), _ptr(col), _ptr(dims), _ptr(code),
                            C.byref(coded)), "xh_csr_grid")
    return (tuple(int(d) for d in dims) if dims[0] else None), code[:len(col)], coded.value


def grid_schedule(nx, ny, nz, nv, prec=0, dots=False):
    """(grid, items per workgroup, [nitems, 4] items: first line y, first line z, x0, x1)"""
    g, ni, ipw = C.c_int64(), C.c_int64(), C.c_int64()
    check(hip().xh_grid_schedule(nx, ny, nz, prec, nv, int(dots), C.byref(g), C.byref(ni),
                                 C.byref(ipw), None), "xh_grid_schedule")
    items = np.zeros((ni.value, 4), np.int32)
    check(hip().xh_grid_schedule(nx, ny, nz, prec, nv, int(dots), None, None, None,
                                 _ptr(items)), "xh_grid_schedule")
    return g.value, ipw.value, items


HBM_PEAK = 8.0e12  # MI355X published HBM3E bandwidth, bytes/s


def kernel_table(stats, peak=HBM_PEAK):
    """rows of a Solver.profile result with per-launch bytes and microseconds, achieved
    algorithmic GB/s and its fraction of `peak` (bytes/s)"""
    rows = []
    for k in stats:
        gbs = k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else 0.0
        rows.append(dict(kernel=k["kernel"], level=k["level"], launches=k["launches"],
                         mb_per_launch=round(k["bytes"] / k["launches"] / 1e6, 3),
                         us_per_launch=round(k["ms"] * 1e3 / k["launches"], 2),
                         ms_total=round(k["ms"], 4), gb_s=round(gbs, 1),
                         frac=round(gbs * 1e9 / peak, 4)))
    return rows


def copy_timed(dst, src, count, reps=20):
    """milliseconds per launch of the library's fp64 copy kernel over count elements"""
    ms = C.c_double()
    check(hip().xh_copy_device_timed(dst.ptr, src.ptr, count, reps, C.byref(ms)),
          "xh_copy_device_timed")
    return ms.value


def set_option(key, value):
    check(hip().xh_set_option(key.encode(), int(value)), "xh_set_option")


def rhs_uniform(seed, n, nv):
    b = np.zeros(n * nv, np.float64)
    check(host().xamg_rhs_uniform(seed, n * nv, _ptr(b)), "xamg_rhs_uniform")
    return b


def amg_default_params():
    p = AmgParams()
    host().xamg_amg_default_params(C.byref(p))
    return p


def amg_strength(row, col, val, theta=0.5, max_row_sum=0.9):
    """strength flags per nonzero of the host AMG setup (uint8, 1 strong)"""
    s = np.zeros(len(col), np.uint8)
    check(host().xamg_amg_strength(len(row) - 1, _ptr(row), _ptr(col), _ptr(val), theta,
                                   max_row_sum, _ptr(s)), "xamg_amg_strength")
    return s


def amg_split(row, col, val, strong, coarsen_type=6, seed=2747):
    """C/F splitting of the host AMG setup: int8, 1 C, -1 F"""
    cf = np.zeros(len(row) - 1, np.int8)
    check(host().xamg_amg_split(len(row) - 1, _ptr(row), _ptr(col), _ptr(val), _ptr(strong),
                                coarsen_type, seed, _ptr(cf)), "xamg_amg_split")
    return cf


def amg_interp(row, col, val, strong, cf, interp_type=0, trunc=0.0, pmax=0):
    """interpolation P of a splitting: (nrows, ncols, row, col, val)"""
    h = C.c_void_p()
    cf = np.ascontiguousarray(cf, np.int8)
    check(host().xamg_amg_interp(len(row) - 1, _ptr(row), _ptr(col), _ptr(val), _ptr(strong),
                                 _ptr(cf), interp_type, trunc, pmax, C.byref(h)),
          "xamg_amg_interp")
    a = Amg.__new__(Amg)
    a.h = h
    a._keep = None
    return a.level(0, 1)


class Amg:
    """Host-side classical AMG hierarchy (libxamg.so)."""

    def __init__(self, row, col, val, params=None, level_keys=None, **kw):
        """level_keys: optional list of dicts, one per level from level 0, each with any
        of the AmgLevelKeys fields (the others take the list-wide values)"""
        p = params or amg_default_params()
        for k, v in kw.items():
            setattr(p, k, v)
        self._lev = None
        if level_keys:
            arr = (AmgLevelKeys * len(level_keys))()
            for l, d in enumerate(level_keys):
                q = arr[l]
                for f in ("strength", "max_row_sum", "trunc_factor", "pmax", "coarsen_type",
                          "interp_type", "agg_interp_type", "agg_trunc_factor", "agg_pmax",
                          "num_paths"):
                    setattr(q, f, d.get(f, getattr(p, f)))
            self._lev = arr
            p.nlev_keys = len(level_keys)
            p.lev_keys = C.cast(arr, C.c_void_p)
        self._keep = (row, col, val)
        self.h = C.c_void_p()
        n = len(row) - 1
        check(host().xamg_amg_setup(n, _ptr(row), _ptr(col), _ptr(val), C.byref(p),
                                    C.byref(self.h)), "xamg_amg_setup")

    @property
    def nlevels(self):
        return host().xamg_amg_nlevels(self.h)

    def level(self, l, which):
        """which: 0 A, 1 P, 2 R -> (nrows, ncols, row, col, val) copies"""
        nr, nc, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        r, c, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(host().xamg_amg_level(self.h, l, which, C.byref(nr), C.byref(nc), C.byref(nnz),
                                    C.byref(r), C.byref(c), C.byref(v)), "xamg_amg_level")
        row = np.ctypeslib.as_array(C.cast(r, C.POINTER(C.c_uint32)), (nr.value + 1,)).copy()
        col = (np.ctypeslib.as_array(C.cast(c, C.POINTER(C.c_uint32)), (nnz.value,)).copy()
               if nnz.value else np.zeros(0, np.uint32))
        val = (np.ctypeslib.as_array(C.cast(v, C.POINTER(C.c_double)), (nnz.value,)).copy()
               if nnz.value else np.zeros(0, np.float64))
        return nr.value, nc.value, row, col, val

    def eigs(self, l):
        a, b = C.c_double(), C.c_double()
        check(host().xamg_amg_eigs(self.h, l, C.byref(a), C.byref(b)), "xamg_amg_eigs")
        return a.value, b.value

    def estimate_eigs(self, iters=10, seed=12345):
        check(host().xamg_amg_estimate_eigs(self.h, iters, seed), "xamg_amg_estimate_eigs")

    def export(self):
        """list of dicts with A, P, R (scipy-free numpy CSR triplets) and eigenvalues"""
        out = []
        for l in range(self.nlevels):
            d = {"A": self.level(l, 0), "eig": self.eigs(l)}
            if l + 1 < self.nlevels:
                d["P"] = self.level(l, 1)
                d["R"] = self.level(l, 2)
            out.append(d)
        return out

    def __del__(self):
        if getattr(self, "h", None) and _host is not None:
            _host.xamg_amg_free(self.h)
            self.h = None


class Hier:
    """Device-side hierarchy object (libxamg_hip.so) built from exported levels."""

    def __init__(self, levels):
        self.h = C.c_void_p()
        check(hip().xh_hier_create(len(levels), C.byref(self.h)), "xh_hier_create")
        self._keep = levels
        for l, d in enumerate(levels):
            n, _, ra, ca, va = d["A"]
            if "P" in d:
                _, nc, rp, cp, vp_ = d["P"]
                _, _, rr, cr, vr = d["R"]
                args = (nc, len(cp), _ptr(rp), _ptr(cp), _ptr(vp_), len(cr), _ptr(rr), _ptr(cr),
                        _ptr(vr))
            else:
                args = (0, 0, None, None, None, 0, None, None, None)
            mn, mx = d["eig"]
            check(hip().xh_hier_set_level(self.h, l, n, len(ca), _ptr(ra), _ptr(ca), _ptr(va),
                                          *args, mn, mx), "xh_hier_set_level")

    def __del__(self):
        if getattr(self, "h", None) and _hip is not None:
            _hip.xh_hier_destroy(self.h)
            self.h = None


class Solver:
    def __init__(self, hier, params, nv):
        self.hier = hier
        self.nv = nv
        self.n = hier._keep[0]["A"][0]
        self.s = C.c_void_p()
        check(hip().xh_solver_create(hier.h, C.byref(params), nv, C.byref(self.s)),
              "xh_solver_create")

    def solve(self, b, x0=None):
        b = np.ascontiguousarray(b, np.float64).reshape(-1)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        check(hip().xh_solver_solve(self.s, _ptr(b), _ptr(x)), "xh_solver_solve")
        return x

    def solve_device(self, db, dx):
        check(hip().xh_solver_solve_device(self.s, db.ptr, dx.ptr), "xh_solver_solve_device")

    def precond_apply(self, db, dx):
        check(hip().xh_solver_precond_apply(self.s, db.ptr, dx.ptr), "xh_solver_precond_apply")

    def smooth(self, db, dx, sweeps=1):
        """dx <- the finest level's pre-smoother applied `sweeps` times from dx"""
        check(hip().xh_solver_smooth(self.s, db.ptr, dx.ptr, sweeps), "xh_solver_smooth")

    def profile(self, db, dx, iters, max_stats=512):
        """per-(kernel, level) launches, algorithmic bytes and event-timed ms of the init
        step plus `iters` iterations from x = dx (xh_solver_profile); list of dicts,
        slowest first"""
        arr = (KernelStat * max_stats)()
        n = C.c_int32()
        check(hip().xh_solver_profile(self.s, db.ptr, dx.ptr, iters, arr, max_stats,
                                      C.byref(n)), "xh_solver_profile")
        return [dict(kernel=k.kernel.decode(), level=k.level, launches=k.launches,
                     bytes=k.bytes, ms=k.ms) for k in arr[: min(n.value, max_stats)]]

    def stats(self):
        it = C.c_int32()
        conv = np.zeros(self.nv, np.int32)
        a = np.zeros(self.nv)
        r = np.zeros(self.nv)
        check(hip().xh_solver_stats(self.s, C.byref(it), _ptr(conv), _ptr(a), _ptr(r)),
              "xh_solver_stats")
        return it.value, conv, a, r

    def history(self, max_checks=1000):
        h = np.zeros(max_checks * self.nv)
        nc = C.c_int32()
        check(hip().xh_solver_history(self.s, _ptr(h), max_checks, C.byref(nc)),
              "xh_solver_history")
        return h[: nc.value * self.nv].reshape(nc.value, self.nv)

    def info(self):
        nl = C.c_int()
        by = C.c_size_t()
        check(hip().xh_solver_info(self.s, C.byref(nl), C.byref(by)), "xh_solver_info")
        return nl.value, by.value

    def __del__(self):
        if getattr(self, "s", None) and _hip is not None:
            _hip.xh_solver_destroy(self.s)
            self.s = None


class Batch:
    """Batched independent systems (include/xamg_hip.h: xh_batch_*).  systems: list of
    (row, col, val); nvs: RHS count per system; lmax: largest eigenvalue bound of
    D^-1/2 A D^-1/2 per system (Chebyshev)."""

    def __init__(self, systems, nvs, params, lmax=None):
        self.ns = len(systems)
        self.keep = [(np.ascontiguousarray(r, np.uint32), np.ascontiguousarray(c, np.uint32),
                      np.ascontiguousarray(v, np.float64)) for r, c, v in systems]
        self.n = np.array([len(r) - 1 for r, _, _ in self.keep], np.int64)
        self.nvs = np.ascontiguousarray(nvs, np.int32)
        rows = (C.c_void_p * self.ns)(*[r.ctypes.data for r, _, _ in self.keep])
        cols = (C.c_void_p * self.ns)(*[c.ctypes.data for _, c, _ in self.keep])
        vals = (C.c_void_p * self.ns)(*[v.ctypes.data for _, _, v in self.keep])
        self.lmax = None if lmax is None else np.ascontiguousarray(lmax, np.float64)
        self.h = C.c_void_p()
        check(hip().xh_batch_create(self.ns, _ptr(self.n), _ptr(self.nvs), rows, cols, vals,
                                    None if self.lmax is None else _ptr(self.lmax),
                                    C.byref(params), C.byref(self.h)), "xh_batch_create")

    def __del__(self):
        if getattr(self, "h", None) and _hip is not None:
            _hip.xh_batch_destroy(self.h)
            self.h = None

    def set_rhs(self, s, b):
        b = np.ascontiguousarray(b, np.float64)
        assert b.size == self.n[s] * self.nvs[s]
        check(hip().xh_batch_set_rhs(self.h, s, _ptr(b)), "xh_batch_set_rhs")

    def solve(self):
        check(hip().xh_batch_solve(self.h), "xh_batch_solve")

    def x(self, s):
        out = np.zeros(int(self.n[s] * self.nvs[s]))
        check(hip().xh_batch_get_x(self.h, s, _ptr(out)), "xh_batch_get_x")
        return out

    def stats(self, s):
        it = C.c_int32()
        nv = int(self.nvs[s])
        conv, a, r = np.zeros(nv, np.int32), np.zeros(nv), np.zeros(nv)
        lat = C.c_double()
        check(hip().xh_batch_stats(self.h, s, C.byref(it), _ptr(conv), _ptr(a), _ptr(r),
                                   C.byref(lat)), "xh_batch_stats")
        return dict(iters=it.value, conv=conv, abs_res=a, rel_res=r, latency_ms=lat.value)

    def layout(self):
        """(tile passes, RHS per lane) per nv class 1, 2, 4, 8, 16"""
        p = np.zeros(5, np.int32)
        w = np.zeros(5, np.int32)
        check(hip().xh_batch_get_layout(self.h, _ptr(p), _ptr(w)), "xh_batch_get_layout")
        return [int(q) for q in p], [int(q) for q in w]

    def set_layout(self, passes, rhs_per_lane):
        p = np.ascontiguousarray(passes, np.int32)
        w = np.ascontiguousarray(rhs_per_lane, np.int32)
        check(hip().xh_batch_set_layout(self.h, _ptr(p), _ptr(w)), "xh_batch_set_layout")

    def set_persistent(self, on=True):
        """solve in one cooperative launch (True) or one launch sequence per iteration"""
        check(hip().xh_batch_set_persistent(self.h, int(on)), "xh_batch_set_persistent")

    def autotune(self, reps=5):
        t = np.zeros(5, np.int32)
        check(hip().xh_batch_autotune(self.h, reps, _ptr(t)), "xh_batch_autotune")
        return t

    def set_tiles(self, tiles):
        t = np.ascontiguousarray(tiles, np.int32)
        check(hip().xh_batch_set_tiles(self.h, _ptr(t)), "xh_batch_set_tiles")

    def spmv_timed(self, reps=20):
        """(µs per launch, algorithmic bytes) of the batched Chebyshev step over every
        system (xh_batch_spmv_timed); call between solves"""
        us, nb = C.c_double(), C.c_double()
        check(hip().xh_batch_spmv_timed(self.h, reps, C.byref(us), C.byref(nb)),
              "xh_batch_spmv_timed")
        return us.value, nb.value


class UNet:
    """Learned initial guess (include/xamg_unet.h): a 3D U-Net on the GPU."""

    def __init__(self, channels, seed=0):
        self.channels = np.ascontiguousarray(channels, np.int32)
        self.h = C.c_void_p()
        check(hip().xh_unet_create(len(self.channels), _ptr(self.channels), seed,
                                   C.byref(self.h)), "xh_unet_create")

    def __del__(self):
        if getattr(self, "h", None) and _hip is not None:
            _hip.xh_unet_destroy(self.h)
            self.h = None

    def nparams(self):
        n = C.c_int64()
        check(hip().xh_unet_nparams(self.h, C.byref(n)), "xh_unet_nparams")
        return n.value

    def get_params(self):
        p = np.zeros(self.nparams(), np.float32)
        check(hip().xh_unet_get_params(self.h, _ptr(p)), "xh_unet_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams()
        check(hip().xh_unet_set_params(self.h, _ptr(p)), "xh_unet_set_params")

    @property
    def output_scale(self):
        o = C.c_float()
        check(hip().xh_unet_get_output_scale(self.h, C.byref(o)), "xh_unet_get_output_scale")
        return o.value

    @output_scale.setter
    def output_scale(self, o):
        check(hip().xh_unet_set_output_scale(self.h, float(o)), "xh_unet_set_output_scale")

    def loss_grad(self, x_in, tgt):
        """x_in, tgt: float32 [n, d, h, w]; returns (loss, flat gradient, y)"""
        x_in = np.ascontiguousarray(x_in, np.float32)
        tgt = np.ascontiguousarray(tgt, np.float32)
        n, d, h, w = x_in.shape
        loss = C.c_float()
        g = np.zeros(self.nparams(), np.float32)
        y = np.zeros(x_in.size, np.float32)
        check(hip().xh_unet_loss_grad(self.h, n, d, h, w, _ptr(x_in), _ptr(tgt), C.byref(loss),
                                      _ptr(g), _ptr(y)), "xh_unet_loss_grad")
        return loss.value, g, y.reshape(x_in.shape)

    def train(self, nx, ny, nz, pairs, nv, crop, batch, steps, lr=1e-3, seed=0):
        """pairs: list of (b, x) DeviceArray multivectors (nx*ny*nz rows, nv columns)"""
        bs = (C.c_void_p * len(pairs))(*[b.ptr for b, _ in pairs])
        xs = (C.c_void_p * len(pairs))(*[x.ptr for _, x in pairs])
        hist = np.zeros(max(steps, 1), np.float32)
        check(hip().xh_unet_train(self.h, nx, ny, nz, len(pairs), bs, xs, nv, crop, batch, steps,
                                  lr, seed, _ptr(hist)), "xh_unet_train")
        return hist[:steps]

    def infer(self, nx, ny, nz, nv, db, dx0):
        check(hip().xh_unet_infer(self.h, nx, ny, nz, nv, db.ptr, dx0.ptr), "xh_unet_infer")


def mfma_selftest(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    d = np.zeros((16, 16), np.float32)
    check(hip().xh_unet_mfma_selftest(_ptr(a), _ptr(b), _ptr(d)), "xh_unet_mfma_selftest")
    return d


class DeviceArray:
    """A device buffer of float64/float32 values (xh_malloc)."""

    def __init__(self, count, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        self.p = C.c_void_p()
        check(hip().xh_malloc(C.byref(self.p), max(self.nbytes, 8)), "xh_malloc")

    @property
    def ptr(self):
        return self.p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        check(hip().xh_memcpy_h2d(d.p, _ptr(a), a.nbytes), "xh_memcpy_h2d")
        return d

    def to_host(self):
        a = np.empty(self.count, self.dtype)
        check(hip().xh_memcpy_d2h(_ptr(a), self.p, self.nbytes), "xh_memcpy_d2h")
        return a

    def zero(self):
        check(hip().xh_fill_zero(self.p, self.nbytes), "xh_fill_zero")

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        check(hip().xh_memcpy_h2d(self.p, _ptr(a), a.nbytes), "xh_memcpy_h2d")

    def __del__(self):
        if getattr(self, "p", None) and _hip is not None:
            _hip.xh_free(self.p)
            self.p = None


class Matrix:
    def __init__(self, row, col, val, prec=0, index_mode=0, ncols=None):
        n = len(row) - 1
        self.row, self.col, self.val = (np.ascontiguousarray(row, np.uint32),
                                        np.ascontiguousarray(col, np.uint32),
                                        np.ascontiguousarray(val, np.float64))
        self.m = C.c_void_p()
        check(hip().xh_matrix_create(n, ncols if ncols is not None else n, len(col),
                                     _ptr(self.row), _ptr(self.col), _ptr(self.val), prec,
                                     index_mode, C.byref(self.m)), "xh_matrix_create")

    @classmethod
    def compressed(cls, y_rows, row, col, val, row_ind=None, col_ind=None, col_offset=0,
                   x_rows=0, prec=0):
        """A compressed block (include/xamg_hip.h: xh_matrix_create_compressed); col keeps
        its dtype (uint8, uint16 or uint32) as the compact-index width."""
        self = cls.__new__(cls)
        self.row = np.ascontiguousarray(row, np.uint32)
        self.col = np.ascontiguousarray(col)
        assert self.col.dtype in (np.uint8, np.uint16, np.uint32)
        self.val = np.ascontiguousarray(val, np.float64)
        self.row_ind = None if row_ind is None else np.ascontiguousarray(row_ind, np.uint32)
        self.col_ind = None if col_ind is None else np.ascontiguousarray(col_ind, np.uint32)
        n = len(self.row) - 1
        ncols = len(self.col_ind) if self.col_ind is not None else int(self.col.max()) + 1
        self.m = C.c_void_p()
        check(hip().xh_matrix_create_compressed(
            y_rows, n, ncols, len(self.col), _ptr(self.row), _ptr(self.col),
            self.col.dtype.itemsize, None if self.row_ind is None else _ptr(self.row_ind),
            None if self.col_ind is None else _ptr(self.col_ind), col_offset, x_rows,
            _ptr(self.val), prec, C.byref(self.m)), "xh_matrix_create_compressed")
        return self

    def gather(self, nv, dx_full, dx_c):
        check(hip().xh_matrix_gather(self.m, nv, dx_full.ptr, dx_c.ptr), "xh_matrix_gather")

    def info(self):
        nr, nc, nnz, pr, bits, panel = (C.c_int64(), C.c_int64(), C.c_int64(), C.c_int(),
                                        C.c_int(), C.c_int64())
        by = C.c_size_t()
        check(hip().xh_matrix_info(self.m, C.byref(nr), C.byref(nc), C.byref(nnz), C.byref(pr),
                                   C.byref(bits), C.byref(panel), C.byref(by)), "xh_matrix_info")
        return dict(nrows=nr.value, ncols=nc.value, nnz=nnz.value, prec=pr.value,
                    index_bits=bits.value, panel_rows=panel.value, bytes=by.value)

    def spmv(self, nv, dx, dy, accumulate=False):
        check(hip().xh_spmv(self.m, nv, dx.ptr, dy.ptr, int(accumulate)), "xh_spmv")

    def fold(self, set=None):
        f = C.c_int64()
        check(hip().xh_matrix_fold(self.m, -1 if set is None else int(set), C.byref(f)),
              "xh_matrix_fold")
        return f.value

    def grid(self):
        """(nx, ny, nz) of the structured-grid view, or None"""
        d = np.zeros(3, np.int32)
        check(hip().xh_matrix_grid(self.m, _ptr(d)), "xh_matrix_grid")
        return tuple(int(v) for v in d) if d[0] else None

    def layouts(self):
        """set of the device layouts the matrix carries ("grid", "st7", "u16")"""
        f = C.c_int32()
        check(hip().xh_matrix_layouts(self.m, C.byref(f)), "xh_matrix_layouts")
        return {n for b, n in ((1, "grid"), (2, "st7"), (4, "u16")) if f.value & b}

    def spmv_timed(self, nv, dx, dy, reps, accumulate=False):
        ms = C.c_double()
        check(hip().xh_spmv_timed(self.m, nv, dx.ptr, dy.ptr, int(accumulate), reps,
                                  C.byref(ms)), "xh_spmv_timed")
        return ms.value

    def __del__(self):
        if getattr(self, "m", None) and _hip is not None:
            _hip.xh_matrix_destroy(self.m)
            self.m = None


def init(device=0):
    check(hip().xh_init(device), "xh_init")


def device_info():
    name = C.create_string_buffer(256)
    cu = C.c_int()
    mem = C.c_double()
    check(hip().xh_device_info(name, 256, C.byref(cu), C.byref(mem)), "xh_device_info")
    return name.value.decode(), cu.value, mem.value


def default_params(method="MergedPBiCGStab", precond="MultiGrid", smoother="Chebyshev",
                   poly_order=4, sweeps=1, rel_tol=1e-8, abs_tol=0.0, max_iters=50,
                   coarse="Direct", mg_prec=0, hsgs_block_rows=64, graph=1, index_mode=0,
                   mg_cycle=0, relax=0.99, fraction=0.3, coarse_fp64=1, true_res_period=0):
    p = Params()
    p.solver = role(method, max_iters=max_iters, abs_tol=abs_tol, rel_tol=rel_tol,
                    conv_check=1)
    # poly_order, relax and fraction apply to every role that uses them: the smoothers
    # and a Chebyshev / Jacobi / HSGS preconditioner without MultiGrid
    sm = dict(max_iters=sweeps, relax_factor=relax, poly_order=poly_order,
              spectrum_fraction=fraction)
    p.precond = (role(precond, max_iters=1, relax_factor=relax, poly_order=poly_order,
                      spectrum_fraction=fraction) if precond else role(NONE))
    p.pre = role(smoother, **sm)
    p.post = role(smoother, **sm)
    p.coarse = role(coarse, **sm)
    p.mg_cycle = mg_cycle
    p.mg_prec = mg_prec
    p.hsgs_ncores = 1
    p.hsgs_block_rows = hsgs_block_rows
    p.coarse_fp64 = coarse_fp64
    p.index_mode = index_mode
    p.graph = graph
    p.true_res_period = true_res_period
    p.nlev_roles = 0
    return p


def set_level_roles(p, pre=None, post=None):
    """Per-level smoother roles: pre[l] / post[l] (Role objects or None for the list-wide
    role) become the roles of level l, the other levels keep p.pre / p.post."""
    pre, post = list(pre or []), list(post or [])
    n = max(len(pre), len(post))
    if n > MAX_LEVELS:
        raise ValueError("at most %d levels" % MAX_LEVELS)
    for l in range(n):
        a = pre[l] if l < len(pre) and pre[l] is not None else p.pre
        b = post[l] if l < len(post) and post[l] is not None else p.post
        p.pre_lev[l] = Role(a.method, a.max_iters, a.abs_tol, a.rel_tol, a.conv_check,
                            a.relax_factor, a.poly_order, a.spectrum_fraction)
        p.post_lev[l] = Role(b.method, b.max_iters, b.abs_tol, b.rel_tol, b.conv_check,
                             b.relax_factor, b.poly_order, b.spectrum_fraction)
    p.nlev_roles = n
    return p
