"""Generators of the variable-coefficient benchmark operators and smooth right-hand sides
(libxamg.so, host only): closed forms, exact symmetry, the reduction to the reference's
cube generator at k = 1."""
import numpy as np
import scipy.sparse as sp


def test_varcoef7_with_unit_k_is_the_cube(X):
    for dims in ((9, 7, 6), (5, 5, 5), (12, 1, 1), (6, 4, 1)):
        row, col, val = X.cube(*dims)
        r2, c2, v2, k = X.varcoef7(*dims, sigma=0.0)
        assert (k == 1.0).all()
        assert np.array_equal(row, r2) and np.array_equal(col, c2)
        assert np.array_equal(val, v2)  # bit for bit, same diagonal accumulation order


def test_varcoef7_is_symmetric_with_zero_interior_row_sums(X):
    N = 10
    row, col, val, k = X.varcoef7(N, sigma=1.0, seed=3)
    A = sp.csr_matrix((val, col.astype(np.int64), row.astype(np.int64)), shape=(N ** 3,) * 2)
    assert (A != A.T).nnz == 0  # exactly symmetric
    rs = np.asarray(A.sum(axis=1)).ravel()
    ii, jj, kk = np.meshgrid(range(N), range(N), range(N), indexing="ij")
    interior = ((ii > 0) & (ii < N - 1) & (jj > 0) & (jj < N - 1) & (kk > 0) & (kk < N - 1))
    interior = interior.transpose(2, 1, 0).ravel()  # x fastest
    scale = np.abs(val).max()
    assert np.abs(rs[interior]).max() < 1e-12 * scale
    assert (rs[~interior] > 0).all()
    # one interior face: coefficient is the harmonic mean times 1/h^2
    p = 5 + N * (5 + N * 5)
    q = p + 1
    j = [t for t in range(row[p], row[p + 1]) if col[t] == q][0]
    assert np.isclose(-val[j], N * N * 2 * k[p] * k[q] / (k[p] + k[q]), rtol=1e-15)


def test_lognormal_moments(X):
    k = X.lognormal_field(11, 2.0, 200000)
    z = np.log(k)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 2.0) < 0.02
    assert np.array_equal(k, X.lognormal_field(11, 2.0, 200000))  # deterministic


def test_q1fem27_constant_k_stencil(X):
    N = 6
    nc = (N + 1) ** 3
    row, col, val, _ = X.q1fem27(N, kcell=np.ones(nc))
    assert len(col) == (3 * N - 2) ** 3
    h = 1.0 / (N + 1)
    A = sp.csr_matrix((val, col.astype(np.int64), row.astype(np.int64)), shape=(N ** 3,) * 2)
    p = 2 + N * (2 + N * 2)  # interior node
    rowp = dict(zip(col[row[p]:row[p + 1]].tolist(), val[row[p]:row[p + 1]].tolist()))
    assert col[row[p]] == p and np.isclose(rowp[p], 8 * h / 3, rtol=1e-14)
    for (di, dj, dl), want in (((1, 0, 0), 0.0), ((0, 1, 1), -h / 6), ((1, -1, 0), -h / 6),
                               ((1, 1, 1), -h / 12), ((-1, 1, -1), -h / 12)):
        q = p + di + N * (dj + N * dl)
        assert np.isclose(rowp[q], want, rtol=1e-14, atol=1e-18)
    assert abs(sum(rowp.values())) < 1e-15  # interior row sum
    assert np.allclose((A - A.T).data, 0.0, atol=1e-15)
    w = np.linalg.eigvalsh(A.toarray())
    assert w.min() > 0


def test_q1fem27_variable_k_is_spd_and_symmetric(X):
    N = 5
    row, col, val, k = X.q1fem27(N, sigma=2.0, seed=4)
    assert len(k) == (N + 1) ** 3
    A = sp.csr_matrix((val, col.astype(np.int64), row.astype(np.int64)), shape=(N ** 3,) * 2)
    assert abs(A - A.T).max() <= 1e-14 * abs(A).max()
    assert np.linalg.eigvalsh(A.toarray()).min() > 0
    # diagonal first, then ascending columns
    for i in range(N ** 3):
        c = col[row[i]:row[i + 1]]
        assert c[0] == i and (np.diff(c[1:].astype(np.int64)) > 0).all()


def test_rhs_bumps_shape_and_determinism(X):
    b = X.rhs_bumps(3, 20, 18, 16, 4)
    assert b.shape == (20 * 18 * 16 * 4,)
    assert np.array_equal(b, X.rhs_bumps(3, 20, 18, 16, 4))
    B = b.reshape(16, 18, 20, 4)
    assert (np.abs(B).max(axis=(0, 1, 2)) > 0).all()
    # smooth: neighbouring cells differ far less than the field's range
    d = np.abs(np.diff(B, axis=2)).max() / np.abs(B).max()
    assert d < 0.5
