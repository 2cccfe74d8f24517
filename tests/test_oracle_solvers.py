"""Pins of the oracle's smoothers, multigrid cycle and Krylov solvers against textbook
matrix-form statements (numpy/scipy), closed forms and convergence theorems."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from conftest import random_csr, to_dense


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def spd_csr(n, seed):
    rng = np.random.default_rng(seed)
    row, col, val = random_csr(rng, n, n, 0.3)
    D = to_dense(row, col, val)
    S = (D + D.T) / 2 + n * np.eye(n)
    A = sp.csr_matrix(S)
    A.sort_indices()
    # diagonal first, then ascending columns
    r, c, v = [0], [], []
    for i in range(n):
        cs = A.indices[A.indptr[i]:A.indptr[i + 1]].tolist()
        vs = A.data[A.indptr[i]:A.indptr[i + 1]].tolist()
        order = sorted(range(len(cs)), key=lambda q: (cs[q] != i, cs[q]))
        c += [cs[q] for q in order]
        v += [vs[q] for q in order]
        r.append(len(c))
    return np.array(r, np.uint32), np.array(c, np.uint32), np.array(v), S


def test_jacobi_matrix_form(O):
    row, col, val, S = spd_csr(20, 1)
    nv, w = 3, np.float32(0.8)
    b = np.random.default_rng(2).standard_normal(20 * nv)
    x = np.zeros(20 * nv)
    A = O.csr(row, col, val)
    O.lib().xo_jacobi_d(C.byref(A), nv, float(w), 3, _p(b), _p(x), 1)
    Dinv = np.diag(1 / np.diag(S))
    B = b.reshape(-1, nv)
    X = np.zeros_like(B)
    for _ in range(3):
        X = X + float(w) * Dinv @ (B - S @ X)
    np.testing.assert_allclose(x.reshape(-1, nv), X, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("nblocks", [1, 3])
def test_hsgs_is_blockwise_ssor(O, nblocks):
    # symmetric SOR sweeps in matrix-splitting form on each core block, with the
    # off-block coupling taken from the previous iterate (block-Jacobi between blocks)
    n, nv, w = 30, 2, np.float32(0.99)
    row, col, val, S = spd_csr(n, 3)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(n * nv)
    x0 = rng.standard_normal(n * nv)
    x = x0.copy()
    A = O.csr(row, col, val)
    O.lib().xo_hsgs_d(C.byref(A), nv, nblocks, float(w), 1, _p(b), _p(x), 0)
    off = np.zeros(nblocks + 1, np.uint64)
    O.lib().xo_core_ranges(n, nblocks, _p(off))
    X = x0.reshape(-1, nv).copy()
    B = b.reshape(-1, nv)
    owner = np.zeros(n, int)
    for c in range(nblocks):
        owner[off[c]:off[c + 1]] = c
    inblk = owner[:, None] == owner[None, :]
    Sin = np.where(inblk, S, 0.0)
    T = B - np.where(inblk, 0.0, S) @ X
    D = np.diag(np.diag(Sin))
    L = np.tril(Sin, -1)
    U = np.triu(Sin, 1)
    wf = float(w)
    X1 = sla.solve_triangular(D / wf + L, (D / wf - D - U) @ X + T, lower=True)
    X2 = sla.solve_triangular(D / wf + U, (D / wf - D - L) @ X1 + T, lower=False)
    np.testing.assert_allclose(x.reshape(-1, nv), X2, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("order", [1, 2, 4])
def test_chebyshev_smoother_matrix_form(O, order):
    n, nv = 25, 2
    row, col, val, S = spd_csr(n, 5)
    rng = np.random.default_rng(6)
    b = rng.standard_normal(n * nv)
    x0 = rng.standard_normal(n * nv)
    d = np.diag(S)
    Ah = S / np.sqrt(d)[:, None] / np.sqrt(d)[None, :]
    ev = np.linalg.eigvalsh(Ah)
    mn, mx, fr = float(ev[0]), float(ev[-1]), np.float32(0.3)
    c = O.cheby_coeffs(mn, mx, order, float(fr))
    P = sum(ci * np.linalg.matrix_power(Ah, i) for i, ci in enumerate(c))
    Dm = np.diag(1 / np.sqrt(d))
    X = x0.reshape(-1, nv) + Dm @ P @ Dm @ (b.reshape(-1, nv) - S @ x0.reshape(-1, nv))
    x = x0.copy()
    A = O.csr(row, col, val)
    O.lib().xo_chebyshev_d(C.byref(A), nv, mn, mx, order, float(fr), _p(b), _p(x), 0)
    np.testing.assert_allclose(x.reshape(-1, nv), X, rtol=1e-10, atol=1e-10)
    # the Chebyshev error polynomial of the exact-bound smoother contracts the error
    # for every eigenvalue above the lower bound


def poisson1d(n):
    r, c, v = [0], [], []
    for i in range(n):
        c.append(i)
        v.append(2.0)
        if i > 0:
            c.append(i - 1)
            v.append(-1.0)
        if i < n - 1:
            c.append(i + 1)
            v.append(-1.0)
        r.append(len(c))
    return np.array(r, np.uint32), np.array(c, np.uint32), np.array(v)


def csr_of(M):
    M = sp.csr_matrix(M)
    M.sort_indices()
    n = M.shape[0]
    r, c, v = [0], [], []
    for i in range(n):
        cs = M.indices[M.indptr[i]:M.indptr[i + 1]].tolist()
        vs = M.data[M.indptr[i]:M.indptr[i + 1]].tolist()
        order = sorted(range(len(cs)), key=lambda q: (cs[q] != i, cs[q]))
        c += [cs[q] for q in order]
        v += [vs[q] for q in order]
        r.append(len(c))
    return (n, M.shape[1], np.array(r, np.uint32), np.array(c, np.uint32), np.array(v))


def two_level_poisson():
    nf, ncs = 15, 7
    row, col, val = poisson1d(nf)
    A = to_dense(row, col, val)
    P = np.zeros((nf, ncs))
    for j in range(ncs):
        i = 2 * j + 1
        P[i, j] = 1.0
        P[i - 1, j] += 0.5
        P[i + 1, j] += 0.5
    R = P.T.copy()
    Ac = R @ A @ P
    levels = [{"A": csr_of(A), "P": csr_of(P), "R": csr_of(R), "eig": (0.0, 2.0)},
              {"A": csr_of(Ac), "eig": (0.0, 2.0)}]
    return levels, A, P, R, Ac


def mg_params(O, pre, post, coarse=None, cycle=0, sweeps=1, relax=0.8, order=2):
    p = O.Params()
    p.solver = O.role(O.PBICGSTAB, max_iters=50, rel_tol=1e-10, conv_check=1)
    p.precond = O.role(O.MULTIGRID, max_iters=1)
    p.pre = O.role(pre, max_iters=sweeps, relax_factor=relax, poly_order=order)
    p.post = O.role(post, max_iters=sweeps, relax_factor=relax, poly_order=order)
    p.coarse = O.role(coarse if coarse is not None else O.DIRECT)
    p.mg_cycle = cycle
    p.hsgs_ncores = 1
    p.dot_ncores = 1
    return p


def test_two_level_vcycle_matrix_form(O):
    levels, A, P, R, Ac = two_level_poisson()
    H = O.Hier(levels)
    nv = 2
    rng = np.random.default_rng(0)
    b = rng.standard_normal(15 * nv)
    x = O.mg_apply(H, mg_params(O, O.JACOBI, O.JACOBI, relax=0.8), nv, b)
    w = float(np.float32(0.8))
    Dinv = np.diag(1 / np.diag(A))
    B = b.reshape(-1, nv)
    X = w * Dinv @ B                       # pre-smoothing from zero
    X = X + P @ np.linalg.solve(Ac, R @ (B - A @ X))
    X = X + w * Dinv @ (B - A @ X)         # post-smoothing
    np.testing.assert_allclose(x.reshape(-1, nv), X, rtol=1e-12, atol=1e-12)


def test_w_and_f_cycles_reduce_to_v_on_two_levels(O):
    levels, A, P, R, Ac = two_level_poisson()
    H = O.Hier(levels)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(15)
    xv = O.mg_apply(H, mg_params(O, O.JACOBI, O.JACOBI, cycle=0), 1, b)
    xw = O.mg_apply(H, mg_params(O, O.JACOBI, O.JACOBI, cycle=1), 1, b)
    np.testing.assert_allclose(xv, xw, rtol=1e-14, atol=1e-14)
    # F cycle skips pre-smoothing: x = S_post(P Ac^-1 R b)
    xf = O.mg_apply(H, mg_params(O, O.JACOBI, O.JACOBI, cycle=2), 1, b)
    w = float(np.float32(0.99))
    w = float(np.float32(0.8))
    X = P @ np.linalg.solve(Ac, R @ b)
    X = X + w * (b - A @ X) / np.diag(A)
    np.testing.assert_allclose(xf, X, rtol=1e-12, atol=1e-12)


def test_vcycle_is_linear(O):
    levels, *_ = two_level_poisson()
    H = O.Hier(levels)
    rng = np.random.default_rng(2)
    b1, b2 = rng.standard_normal(15), rng.standard_normal(15)
    prm = mg_params(O, O.HSGS, O.HSGS, relax=0.9)
    m = lambda b: O.mg_apply(H, prm, 1, b)  # noqa: E731
    np.testing.assert_allclose(m(b1 + 2 * b2), m(b1) + 2 * m(b2), rtol=1e-12, atol=1e-12)


def diag_hier(d):
    n = len(d)
    row = np.arange(n + 1, dtype=np.uint32)
    col = np.arange(n, dtype=np.uint32)
    return [{"A": (n, n, row, col, np.asarray(d, float)), "eig": (0.0, 1.0)}]


def krylov_params(O, method, precond=None, tol=1e-12, iters=100):
    p = O.Params()
    p.solver = O.role(method, max_iters=iters, rel_tol=tol, abs_tol=0.0, conv_check=1)
    p.precond = O.role(precond if precond is not None else O.NONE)
    p.pre = p.post = p.coarse = O.role(O.NONE)
    p.hsgs_ncores = 1
    p.dot_ncores = 1
    return p


@pytest.mark.parametrize("method", [6, 7, 8, 9])
def test_scaled_identity_breakdown(O, method):
    # A = c I: alpha = 1/c exactly, so s = r - alpha v = 0, t = 0 and omega = 0/0.
    # The reference has no breakdown guard: res = theta - omega*phi is NaN, the test
    # `abs_res > tol` is false, every RHS reports converged after one iteration and x
    # (updated with omega*s = NaN) is NaN.  The MergedBiCGStab variant updates x inside
    # the converged branch too (merged_bicgstab.inl:166-170).
    H = O.Hier(diag_hier(np.full(40, 2.0)))
    b = np.random.default_rng(3).standard_normal(40 * 2)
    prec = O.IDENTITY if method in (7, 9) else None
    r = O.solve(H, krylov_params(O, method, prec), 2, b)
    assert r["iters"] == 1 and r["conv"].all()
    assert np.isnan(r["x"]).all() and np.isnan(r["abs_res"]).all()


@pytest.mark.parametrize("method", [6, 7, 8, 9])
def test_distinct_eigenvalues_bound(O, method):
    # BiCGStab on a diagonal matrix with m distinct eigenvalues terminates in <= m
    # iterations in exact arithmetic (each iteration raises the Krylov degree by 2)
    d = np.repeat([1.0, 2.0, 5.0, 11.0], 25)
    H = O.Hier(diag_hier(d))
    b = np.random.default_rng(4).standard_normal(100 * 3)
    prec = O.IDENTITY if method in (7, 9) else None
    r = O.solve(H, krylov_params(O, method, prec, tol=1e-10), 3, b)
    assert r["conv"].all() and r["iters"] <= 4
    np.testing.assert_allclose(r["x"].reshape(-1, 3), b.reshape(-1, 3) / d[:, None], rtol=1e-8)


def test_pbicgstab_and_merged_are_the_same_arithmetic(O, X):
    row, col, val = O.cube(12)
    levels = X.Amg(row, col, val).export()
    H = O.Hier(levels)
    b = X.rhs_uniform(0, 12 ** 3, 4)
    out = []
    for m in (O.PBICGSTAB, O.MERGED_PBICGSTAB):
        p = mg_params(O, O.CHEBYSHEV, O.CHEBYSHEV, order=3)
        p.solver = O.role(m, max_iters=50, rel_tol=1e-8, conv_check=1)
        out.append(O.solve(H, p, 4, b))
    assert out[0]["iters"] == out[1]["iters"]
    assert np.array_equal(out[0]["x"], out[1]["x"])


def test_cube_solution_matches_direct_solve(O, X):
    N, nv = 16, 3
    row, col, val = O.cube(N)
    levels = X.Amg(row, col, val).export()
    H = O.Hier(levels)
    b = X.rhs_uniform(5, N ** 3, nv)
    p = mg_params(O, O.HSGS, O.HSGS, relax=0.99)
    p.solver = O.role(O.PBICGSTAB, max_iters=50, rel_tol=1e-10, conv_check=1)
    r = O.solve(H, p, nv, b, hist_len=60)
    A = sp.csr_matrix((val, col, row), shape=(N ** 3, N ** 3))
    xs = spla.spsolve(A.tocsc(), b.reshape(-1, nv))
    assert r["conv"].all()
    err = np.abs(r["x"].reshape(-1, nv) - xs).max(0) / np.abs(xs).max(0)
    assert err.max() < 1e-8
    # the tested residual is the recurrence residual theta - omega*phi = ||s - w t||^2;
    # at convergence it agrees with the true residual of the returned x
    tr = np.linalg.norm(b.reshape(-1, nv) - A @ r["x"].reshape(-1, nv), axis=0)
    np.testing.assert_allclose(r["abs_res"], tr, rtol=1e-4)
    # residual history is monotone enough to reach 1e-10 and starts at ||b||
    np.testing.assert_allclose(r["hist"][0], np.linalg.norm(b.reshape(-1, nv), axis=0),
                               rtol=1e-14)


def test_uneven_convergence_freezes_converged_rhs(O, X):
    # RHS k is scaled by 10^-k; with an absolute tolerance the small ones converge
    # first and their x must stop changing (iconv masking, merged_pbicgstab.inl:191-199)
    N, nv = 12, 4
    row, col, val = O.cube(N)
    levels = X.Amg(row, col, val).export()
    H = O.Hier(levels)
    b = X.rhs_uniform(1, N ** 3, nv).reshape(-1, nv) * (10.0 ** -np.arange(nv) * 3)
    b = b.reshape(-1)
    p = mg_params(O, O.JACOBI, O.JACOBI, relax=0.7, sweeps=2)
    p.solver = O.role(O.PBICGSTAB, max_iters=60, rel_tol=0.0, abs_tol=1e-6, conv_check=1)
    r = O.solve(H, p, nv, b, hist_len=80)
    h = r["hist"]
    # first check index at which each RHS meets the absolute tolerance
    first = [int(np.argmax(h[:, k] <= np.float32(1e-6))) for k in range(nv)]
    assert first == sorted(first, reverse=True) and first[0] > first[-1]
    assert r["conv"].all()
    # a converged RHS keeps the abs residual recorded at its first convergence
    for k in range(nv):
        assert r["abs_res"][k] == pytest.approx(h[first[k], k], rel=1e-12)


@pytest.mark.parametrize("method", ["PBICGSTAB", "MERGED_PBICGSTAB", "BICGSTAB"])
def test_true_residual_replacement(O, X, method):
    # r <- b - A x every k-th iteration: in exact arithmetic the iterates do not change,
    # so on a well-conditioned operator x and the iteration count stay put, and with a
    # replacement every iteration the tested (recurrence) residual follows the true
    # residual of the returned x closely
    N, nv = 14, 3
    row, col, val = O.cube(N)
    levels = X.Amg(row, col, val).export()
    H = O.Hier(levels)
    b = X.rhs_uniform(8, N ** 3, nv)
    A = sp.csr_matrix((val, col, row), shape=(N ** 3, N ** 3))
    m = getattr(O, method)
    res = {}
    for k in (0, 1, 3):
        p = mg_params(O, O.CHEBYSHEV, O.CHEBYSHEV, order=2)
        if method == "BICGSTAB":
            p.precond = O.role(O.NONE)
        p.solver = O.role(m, max_iters=300, rel_tol=1e-10, conv_check=1)
        p.true_res_period = k
        res[k] = O.solve(H, p, nv, b)
        assert res[k]["conv"].all()
    x0 = res[0]["x"].reshape(-1, nv)
    for k in (1, 3):
        assert abs(res[k]["iters"] - res[0]["iters"]) <= 1
        err = np.abs(res[k]["x"].reshape(-1, nv) - x0).max(0) / np.abs(x0).max(0)
        assert err.max() < 1e-8
    tr = np.linalg.norm(b.reshape(-1, nv) - A @ res[1]["x"].reshape(-1, nv), axis=0)
    np.testing.assert_allclose(res[1]["abs_res"], tr, rtol=1e-3)


def test_true_residual_replacement_rejected_for_merged_bicgstab(O, X):
    row, col, val = O.cube(8)
    levels = X.Amg(row, col, val).export()
    p = mg_params(O, O.JACOBI, O.JACOBI)
    p.precond = O.role(O.NONE)
    p.solver = O.role(O.MERGED_BICGSTAB, max_iters=50, rel_tol=1e-8, conv_check=1)
    p.true_res_period = 2
    with pytest.raises(RuntimeError):
        O.solve(O.Hier(levels), p, 1, X.rhs_uniform(0, 8 ** 3, 1))


# ---- PCG, RBiCGStab, MergedRBiCGStab (pcg.inl, rbicgstab.inl, merged_rbicgstab.inl) ----

def _krylov_basis(S, Minv, b, k):
    """orthonormal basis of K_k(M^-1 S, M^-1 b)"""
    v = Minv @ b
    cols = []
    for _ in range(k):
        cols.append(v)
        v = Minv @ (S @ v)
    Q, _ = np.linalg.qr(np.array(cols).T)
    return Q


@pytest.mark.parametrize("precond", ["identity", "jacobi"])
def test_pcg_iterate_is_the_krylov_energy_minimiser(O, precond):
    # CG with an SPD preconditioner M: x_k minimises the S-norm error over
    # x0 + K_k(M^-1 S, M^-1 r0); with x0 = 0 that is Q (Q^T S Q)^-1 Q^T b
    n, nv, w = 30, 2, np.float32(0.7)
    row, col, val, S = spd_csr(n, 11)
    levels = [{"A": (n, n, row, col, val), "eig": (0.0, 1.0)}]
    H = O.Hier(levels)
    b = np.random.default_rng(12).standard_normal(n * nv)
    B = b.reshape(-1, nv)
    Minv = np.eye(n) if precond == "identity" else float(w) * np.diag(1 / np.diag(S))
    for k in (1, 2, 4, 6):
        p = krylov_params(O, O.PCG, O.IDENTITY if precond == "identity" else O.JACOBI,
                          tol=0.0, iters=k)
        p.precond.relax_factor = w
        r = O.solve(H, p, nv, b)
        assert r["iters"] == k
        X = r["x"].reshape(-1, nv)
        for c in range(nv):
            Q = _krylov_basis(S, Minv, B[:, c], k)
            xk = Q @ np.linalg.solve(Q.T @ S @ Q, Q.T @ B[:, c])
            np.testing.assert_allclose(X[:, c], xk, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("method", [10, 11, 12])
def test_distinct_eigenvalues_bound_new_methods(O, method):
    d = np.repeat([1.0, 2.0, 5.0, 11.0], 25)
    H = O.Hier(diag_hier(d))
    b = np.random.default_rng(4).standard_normal(100 * 3)
    r = O.solve(H, krylov_params(O, method, O.IDENTITY, tol=1e-10), 3, b)
    assert r["conv"].all() and r["iters"] <= 4
    np.testing.assert_allclose(r["x"].reshape(-1, 3), b.reshape(-1, 3) / d[:, None], rtol=1e-8)


def test_scaled_identity_new_methods(O):
    # PCG on c I: alpha = 1/c, r = 0 after one step, x = b / c exactly
    H = O.Hier(diag_hier(np.full(40, 2.0)))
    b = np.random.default_rng(3).standard_normal(40 * 2)
    r = O.solve(H, krylov_params(O, O.PCG, O.IDENTITY), 2, b)
    assert r["iters"] == 1 and r["conv"].all()
    np.testing.assert_array_equal(r["x"], b / 2.0)
    # RBiCGStab: t_ = z - alpha s = 0, so omega = theta / phi = 0/0 as in BiCGStab
    for m in (O.RBICGSTAB, O.MERGED_RBICGSTAB):
        r = O.solve(H, krylov_params(O, m, O.IDENTITY), 2, b)
        assert r["iters"] == 1 and r["conv"].all() and np.isnan(r["x"]).all()


def test_rbicgstab_is_pbicgstab_reformulated(O, X):
    # with a linear preconditioner RBiCGStab computes the iterates of right-preconditioned
    # BiCGStab (v_ = M^-1 p, t_ = M^-1 s, r^ = s): equal residual histories up to rounding
    row, col, val = O.cube(12)
    levels = X.Amg(row, col, val).export()
    H = O.Hier(levels)
    b = X.rhs_uniform(0, 12 ** 3, 4)
    out = {}
    for m in (O.PBICGSTAB, O.RBICGSTAB, O.MERGED_RBICGSTAB):
        p = mg_params(O, O.CHEBYSHEV, O.CHEBYSHEV, order=3)
        p.solver = O.role(m, max_iters=50, rel_tol=1e-8, conv_check=1)
        out[m] = O.solve(H, p, 4, b, hist_len=60)
    ref = out[O.PBICGSTAB]
    for m in (O.RBICGSTAB, O.MERGED_RBICGSTAB):
        r = out[m]
        assert r["conv"].all() and abs(r["iters"] - ref["iters"]) <= 1
        k = min(5, len(r["hist"]), len(ref["hist"]))
        np.testing.assert_allclose(r["hist"][:k], ref["hist"][:k], rtol=1e-8,
                                   atol=1e-12 * ref["hist"][0].max())
        # both stop at rel-res 1e-8, possibly one iteration apart
        np.testing.assert_allclose(r["x"], ref["x"], rtol=0, atol=1e-4 * np.abs(ref["x"]).max())
    # the merged and plain forms differ only in the association of the v_ update (and the
    # x update of a RHS in the iteration it converges): not bit-identical, same path
    a, c = out[O.RBICGSTAB], out[O.MERGED_RBICGSTAB]
    np.testing.assert_allclose(a["hist"][:4], c["hist"][:4], rtol=1e-10,
                               atol=1e-13 * a["hist"][0].max())
    assert not np.array_equal(a["x"], c["x"])


def test_pcg_solves_cube_with_amg(O, X):
    N, nv = 12, 2
    row, col, val = O.cube(N)
    levels = X.Amg(row, col, val).export()
    H = O.Hier(levels)
    b = X.rhs_uniform(2, N ** 3, nv)
    p = mg_params(O, O.JACOBI, O.JACOBI, relax=0.8, sweeps=2)
    p.solver = O.role(O.PCG, max_iters=60, rel_tol=1e-10, conv_check=1)
    r = O.solve(H, p, nv, b)
    A = sp.csr_matrix((val, col, row), shape=(N ** 3, N ** 3))
    xs = spla.spsolve(A.tocsc(), b.reshape(-1, nv))
    assert r["conv"].all()
    assert (np.abs(r["x"].reshape(-1, nv) - xs).max(0) / np.abs(xs).max(0)).max() < 1e-8


def test_true_residual_rejected_for_new_methods(O, X):
    row, col, val = O.cube(8)
    levels = X.Amg(row, col, val).export()
    for m in (O.PCG, O.RBICGSTAB, O.MERGED_RBICGSTAB):
        p = mg_params(O, O.JACOBI, O.JACOBI)
        p.solver = O.role(m, max_iters=50, rel_tol=1e-8, conv_check=1)
        p.true_res_period = 5
        with pytest.raises(RuntimeError):
            O.solve(O.Hier(levels), p, 1, X.rhs_uniform(0, 8 ** 3, 1))
