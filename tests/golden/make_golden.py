"""Writes the golden vectors of tests/golden/ (run: python tests/golden/make_golden.py).

Independent of the oracle and of the library: the operator is assembled here from the
stencil definition of the reference generator (examples/common/generator/cube.inl:84-146,
restated in numpy), the right-hand sides come from a numpy SplitMix64, and the exact
solutions from scipy.sparse.linalg.spsolve (a direct LU solve).
"""
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def splitmix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def rhs_uniform(seed, count):
    out = np.empty(count)
    for i in range(count):
        z = splitmix64((seed * 0xD1B54A32D192ED03 + i) & M64)
        out[i] = 2.0 * ((z >> 11) * (1.0 / 9007199254740992.0)) - 1.0
    return out


def cube_dense_stencil(nx, ny, nz):
    """7-point operator with ghost-cell Dirichlet boundaries as a scipy matrix"""
    n = nx * ny * nz
    q = (nx * nx, ny * ny, nz * nz)
    idx = np.arange(n)
    ii, jj, kk = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    for (coord, d, qq, stride) in ((ii, nx, q[0], 1), (jj, ny, q[1], nx), (kk, nz, q[2], nx * ny)):
        if d == 1:
            continue
        for side in (-1, 1):
            has = (coord > 0) if side < 0 else (coord < d - 1)
            diag += np.where(has, qq, 2 * qq)
            r = idx[has]
            rows.append(r)
            cols.append(r + side * stride)
            vals.append(np.full(len(r), -qq, float))
    rows.append(idx)
    cols.append(idx)
    vals.append(diag)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def main():
    nx, ny, nz, nv, seed = 9, 7, 6, 3, 11
    A = cube_dense_stencil(nx, ny, nz)
    b = rhs_uniform(seed, nx * ny * nz * nv)
    x = spla.spsolve(A.tocsc(), b.reshape(-1, nv))
    np.savez_compressed(os.path.join(HERE, "cube_9x7x6_nv3.npz"), dims=np.array([nx, ny, nz]),
                        nv=nv, seed=seed, indptr=A.indptr.astype(np.int64),
                        indices=A.indices.astype(np.int64), data=A.data, b=b,
                        x_exact=np.asarray(x).reshape(-1))
    # RHS generator golden values for a larger index range
    np.savez_compressed(os.path.join(HERE, "rhs_splitmix_seed1.npz"),
                        seed=1, first=rhs_uniform(1, 4096))


if __name__ == "__main__":
    main()
