"""ctypes binding of the oracle (xamg_oracle.h).  TEST INFRASTRUCTURE ONLY: imported by
tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py, as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libxamg_oracle.so")

IDENTITY, JACOBI, HSGS, CHEBYSHEV, DIRECT, MULTIGRID = 0, 1, 2, 3, 4, 5
BICGSTAB, PBICGSTAB, MERGED_BICGSTAB, MERGED_PBICGSTAB = 6, 7, 8, 9
PCG, RBICGSTAB, MERGED_RBICGSTAB = 10, 11, 12
NONE = -1


class Csr(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64),
                ("row", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


MAX_LEVELS = 32  # XO_MAX_LEVELS


class Role(C.Structure):
    _fields_ = [("method", C.c_int32), ("max_iters", C.c_int32), ("abs_tol", C.c_float),
                ("rel_tol", C.c_float), ("conv_check", C.c_int32), ("relax_factor", C.c_float),
                ("poly_order", C.c_int32), ("spectrum_fraction", C.c_float)]


class Params(C.Structure):
    _fields_ = [("solver", Role), ("precond", Role), ("pre", Role), ("post", Role),
                ("coarse", Role), ("mg_cycle", C.c_int32), ("mg_prec", C.c_int32),
                ("hsgs_ncores", C.c_int32), ("hsgs_block_rows", C.c_int32),
                ("dot_ncores", C.c_int32), ("coarse_fp64", C.c_int32),
                ("true_res_period", C.c_int32), ("nlev_roles", C.c_int32),
                ("pre_lev", Role * MAX_LEVELS), ("post_lev", Role * MAX_LEVELS)]


class HierS(C.Structure):
    _fields_ = [("nlevels", C.c_int32), ("A", C.POINTER(Csr)), ("P", C.POINTER(Csr)),
                ("R", C.POINTER(Csr)), ("min_eig", C.POINTER(C.c_double)),
                ("max_eig", C.POINTER(C.c_double))]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        vp = C.c_void_p
        P = C.POINTER
        _lib.xo_spmv_d.argtypes = [P(Csr), C.c_int, vp, vp, C.c_int]
        _lib.xo_spmv_f.argtypes = [P(Csr), C.c_int, vp, vp, C.c_int]
        _lib.xo_dense_matvec_d.argtypes = [C.c_int64, C.c_int64, vp, C.c_int, vp, vp]
        _lib.xo_dense_matvec_f.argtypes = [C.c_int64, C.c_int64, vp, C.c_int, vp, vp]
        _lib.xo_inverse_d.argtypes = [C.c_int64, vp, vp]
        _lib.xo_inverse_f.argtypes = [C.c_int64, vp, vp]
        _lib.xo_cheby_coeffs.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, vp]
        _lib.xo_cheby_coeffs.restype = C.c_int
        _lib.xo_split_range.argtypes = [C.c_uint64, C.c_int]
        _lib.xo_split_range.restype = C.c_uint64
        _lib.xo_core_ranges.argtypes = [C.c_uint64, C.c_int, vp]
        _lib.xo_hsgs_d.argtypes = [P(Csr), C.c_int, C.c_int, C.c_float, C.c_int, vp, vp, C.c_int]
        _lib.xo_chebyshev_d.argtypes = [P(Csr), C.c_int, C.c_double, C.c_double, C.c_int,
                                        C.c_float, vp, vp, C.c_int]
        _lib.xo_jacobi_d.argtypes = [P(Csr), C.c_int, C.c_float, C.c_int, vp, vp, C.c_int]
        _lib.xo_compress.argtypes = [P(Csr), C.c_int64, vp, vp, vp, vp, P(C.c_int64),
                                     P(C.c_int64)]
        _lib.xo_compress.restype = C.c_int
        _lib.xo_define_int_type.argtypes = [C.c_uint64]
        _lib.xo_define_int_type.restype = C.c_int
        _lib.xo_cube_nnz.argtypes = [C.c_int64] * 3
        _lib.xo_cube_nnz.restype = C.c_int64
        _lib.xo_cube.argtypes = [C.c_int64] * 3 + [vp, vp, vp]
        _lib.xo_solve.argtypes = [P(HierS), P(Params), C.c_int, vp, vp, P(C.c_int32), vp, vp, vp]
        _lib.xo_solve.restype = C.c_int
        _lib.xo_solve_hist.argtypes = [P(HierS), P(Params), C.c_int, vp, vp, P(C.c_int32), vp,
                                       vp, vp, vp, C.c_int32]
        _lib.xo_solve_hist.restype = C.c_int
        _lib.xo_mg_apply.argtypes = [P(HierS), P(Params), C.c_int, vp, vp]
        _lib.xo_mg_apply.restype = C.c_int
        _lib.xo_solver_create.argtypes = [P(HierS), P(Params), P(vp)]
        _lib.xo_solver_create.restype = C.c_int
        _lib.xo_solver_solve.argtypes = [vp, C.c_int, vp, vp, P(C.c_int32), vp, vp, vp]
        _lib.xo_solver_solve.restype = C.c_int
        _lib.xo_solver_free.argtypes = [vp]
        _lib.xo_solver_free.restype = None
        _lib.xo_set_threads.argtypes = [C.c_int]
        _lib.xo_set_threads.restype = C.c_int
        _lib.xo_get_threads.argtypes = []
        _lib.xo_get_threads.restype = C.c_int
    return _lib


def set_threads(n):
    """OpenMP threads of the oracle's loops"""
    if lib().xo_set_threads(int(n)):
        raise ValueError("thread count must be >= 1")


def threads():
    return lib().xo_get_threads()


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def csr(row, col, val, ncols=None):
    row = np.ascontiguousarray(row, np.uint32)
    col = np.ascontiguousarray(col, np.uint32)
    val = np.ascontiguousarray(val, np.float64)
    n = len(row) - 1
    c = Csr(n, n if ncols is None else ncols, len(col), _p(row), _p(col), _p(val))
    c._keep = (row, col, val)
    return c


def spmv(row, col, val, x, nv, y=None, ncols=None, f32=False):
    c = csr(row, col, val, ncols)
    dt = np.float32 if f32 else np.float64
    x = np.ascontiguousarray(x, dt)
    acc = y is not None
    y = np.zeros((len(row) - 1) * nv, dt) if y is None else np.ascontiguousarray(y, dt).copy()
    f = lib().xo_spmv_f if f32 else lib().xo_spmv_d
    f(C.byref(c), nv, _p(x), _p(y), int(acc))
    return y


def compress(row, col, val, ncols):
    """The reference's compress() of a CSR block: dict(row, col (compact ids, uint32),
    row_ind, col_ind, nrows, ncols, width) with width the define_int_type code of ncols."""
    c = csr(row, col, val, ncols)
    n = len(row) - 1
    r_out = np.zeros(n + 1, np.uint32)
    c_out = np.zeros(len(col), np.uint32)
    r_ind = np.zeros(max(n, 1), np.uint32)
    c_ind = np.zeros(max(ncols, 1), np.uint32)
    nr, nc = C.c_int64(), C.c_int64()
    rc = lib().xo_compress(C.byref(c), ncols, _p(r_out), _p(c_out), _p(r_ind), _p(c_ind),
                           C.byref(nr), C.byref(nc))
    if rc:
        raise RuntimeError(f"oracle compress failed with {rc}")
    nr, nc = nr.value, nc.value
    return dict(row=r_out[:nr + 1].copy(), col=c_out, row_ind=r_ind[:nr].copy(),
                col_ind=c_ind[:nc].copy(), nrows=nr, ncols=nc,
                width=lib().xo_define_int_type(nc))


def cheby_coeffs(mn, mx, order, fraction):
    c = np.zeros(order)
    rc = lib().xo_cheby_coeffs(mn, mx, order, fraction, _p(c))
    if rc:
        raise ValueError("order out of range")
    return c


def cube(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    nnz = lib().xo_cube_nnz(nx, ny, nz)
    n = nx * ny * nz
    row = np.zeros(n + 1, np.uint32)
    col = np.zeros(nnz, np.uint32)
    val = np.zeros(nnz)
    lib().xo_cube(nx, ny, nz, _p(row), _p(col), _p(val))
    return row, col, val


class Hier:
    """levels: list of dicts {"A": (n, nc, row, col, val), "P": ..., "R": ..., "eig": (mn, mx)}"""

    def __init__(self, levels):
        nl = len(levels)
        self.A = (Csr * nl)()
        self.P = (Csr * max(1, nl - 1))()
        self.R = (Csr * max(1, nl - 1))()
        self.mn = (C.c_double * nl)()
        self.mx = (C.c_double * nl)()
        keep = []
        for l, d in enumerate(levels):
            n, nc, r, c, v = d["A"]
            m = csr(r, c, v, nc)
            keep.append(m._keep)
            self.A[l] = m
            if l + 1 < nl:
                n, nc, r, c, v = d["P"]
                m = csr(r, c, v, nc)
                keep.append(m._keep)
                self.P[l] = m
                n, nc, r, c, v = d["R"]
                m = csr(r, c, v, nc)
                keep.append(m._keep)
                self.R[l] = m
            self.mn[l], self.mx[l] = d["eig"]
        self._keep = keep
        self.s = HierS(nl, self.A, self.P, self.R, self.mn, self.mx)


def role(method, max_iters=1, abs_tol=0.0, rel_tol=0.0, conv_check=0, relax_factor=0.99,
         poly_order=2, spectrum_fraction=0.3):
    return Role(method, max_iters, abs_tol, rel_tol, conv_check, relax_factor, poly_order,
                spectrum_fraction)


def params_from(xp, dot_ncores=1):
    """oracle parameters from a pyxamg.Params (same field meaning)"""
    p = Params()
    for f in ("solver", "precond", "pre", "post", "coarse"):
        r = getattr(xp, f)
        setattr(p, f, Role(r.method, r.max_iters, r.abs_tol, r.rel_tol, r.conv_check,
                           r.relax_factor, r.poly_order, r.spectrum_fraction))
    p.mg_cycle = xp.mg_cycle
    p.mg_prec = xp.mg_prec
    p.hsgs_ncores = xp.hsgs_ncores
    p.hsgs_block_rows = xp.hsgs_block_rows
    p.dot_ncores = dot_ncores
    p.coarse_fp64 = xp.coarse_fp64
    p.true_res_period = xp.true_res_period
    p.nlev_roles = xp.nlev_roles
    for l in range(xp.nlev_roles):
        for f in ("pre_lev", "post_lev"):
            r = getattr(xp, f)[l]
            getattr(p, f)[l] = Role(r.method, r.max_iters, r.abs_tol, r.rel_tol, r.conv_check,
                                    r.relax_factor, r.poly_order, r.spectrum_fraction)
    return p


def solve(hier, params, nv, b, x0=None, hist_len=0):
    n = hier.A[0].nrows
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n * nv) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    it = C.c_int32()
    conv = np.zeros(nv, np.int32)
    a = np.zeros(nv)
    r = np.zeros(nv)
    if hist_len:
        h = np.zeros(hist_len * nv)
        rc = lib().xo_solve_hist(C.byref(hier.s), C.byref(params), nv, _p(b), _p(x),
                                 C.byref(it), _p(conv), _p(a), _p(r), _p(h), hist_len)
    else:
        h = None
        rc = lib().xo_solve(C.byref(hier.s), C.byref(params), nv, _p(b), _p(x), C.byref(it),
                            _p(conv), _p(a), _p(r))
    if rc:
        raise RuntimeError(f"oracle solve failed with {rc}")
    out = dict(x=x, iters=it.value, conv=conv, abs_res=a, rel_res=r)
    if h is not None:
        out["hist"] = h[: (it.value + 1) * nv].reshape(-1, nv)
    return out


def mg_apply(hier, params, nv, b):
    n = hier.A[0].nrows
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n * nv)
    rc = lib().xo_mg_apply(C.byref(hier.s), C.byref(params), nv, _p(b), _p(x))
    if rc:
        raise RuntimeError(f"oracle mg_apply failed with {rc}")
    return x


class SolverObj:
    """oracle solver with setup separated from the solve (for timing the solve alone)"""

    def __init__(self, hier, params):
        self.hier = hier
        self.params = params
        self.s = C.c_void_p()
        rc = lib().xo_solver_create(C.byref(hier.s), C.byref(params), C.byref(self.s))
        if rc:
            raise RuntimeError(f"oracle solver create failed with {rc}")

    def solve(self, nv, b, x0=None):
        n = self.hier.A[0].nrows
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n * nv) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        it = C.c_int32()
        conv = np.zeros(nv, np.int32)
        a = np.zeros(nv)
        r = np.zeros(nv)
        rc = lib().xo_solver_solve(self.s, nv, _p(b), _p(x), C.byref(it), _p(conv), _p(a), _p(r))
        if rc:
            raise RuntimeError(f"oracle solve failed with {rc}")
        return dict(x=x, iters=it.value, conv=conv, abs_res=a, rel_res=r)

    def __del__(self):
        if getattr(self, "s", None) and _lib is not None:
            _lib.xo_solver_free(self.s)
            self.s = None
