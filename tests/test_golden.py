"""Golden vectors (tests/golden/make_golden.py: numpy/scipy, independent of the oracle and
the library) and the reference's own YAML parameter fixtures."""
import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp

import yaml_params

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_cube():
    g = np.load(os.path.join(GOLD, "cube_9x7x6_nv3.npz"), allow_pickle=False)
    n = int(np.prod(g["dims"]))
    A = sp.csr_matrix((g["data"], g["indices"], g["indptr"]), shape=(n, n))
    return g, A


def test_cube_generators_match_golden_operator(O, X):
    g, A = load_cube()
    nx, ny, nz = (int(v) for v in g["dims"])
    for gen in (O.cube, X.cube):
        row, col, val = gen(nx, ny, nz)
        B = sp.csr_matrix((val, col, row), shape=A.shape)
        assert (abs(B - A) > 0).nnz == 0


def test_rhs_generator_matches_golden(X):
    g = np.load(os.path.join(GOLD, "rhs_splitmix_seed1.npz"), allow_pickle=False)
    assert np.array_equal(X.rhs_uniform(int(g["seed"]), 1, len(g["first"])), g["first"])
    g2, _ = load_cube()
    n = int(np.prod(g2["dims"]))
    assert np.array_equal(X.rhs_uniform(int(g2["seed"]), n, int(g2["nv"])), g2["b"])


@pytest.mark.parametrize("smoother", ["HSGS", "Chebyshev", "Jacobi"])
def test_oracle_solve_matches_direct_golden(O, X, smoother):
    g, A = load_cube()
    nx, ny, nz = (int(v) for v in g["dims"])
    nv = int(g["nv"])
    row, col, val = O.cube(nx, ny, nz)
    levels = X.Amg(row, col, val, coarse_size=20).export()
    p = X.default_params(method="PBiCGStab", smoother=smoother, poly_order=3,
                         sweeps=1 if smoother == "HSGS" else 2, rel_tol=1e-11)
    r = O.solve(O.Hier(levels), O.params_from(p), nv, g["b"])
    assert r["conv"].all()
    xe = g["x_exact"].reshape(-1, nv)
    err = np.abs(r["x"].reshape(-1, nv) - xe).max(0) / np.abs(xe).max(0)
    assert err.max() < 1e-9


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "yaml", "*.yaml"))))
def test_reference_yaml_configs_parse(X, path):
    roles = yaml_params.load(path)
    p = yaml_params.to_params(X, roles)
    assert p.solver.method == X.PBICGSTAB and p.solver.max_iters == 20
    assert p.solver.rel_tol == 0.0 and p.solver.abs_tol == 0.0
    assert p.precond.method == X.MULTIGRID and p.coarse.method == X.DIRECT
    assert p.pre.method in (X.CHEBYSHEV, X.HSGS, X.JACOBI)
    if "Ch-rp" in path:
        assert p.mg_prec == 1
    if "Jac" in path:
        assert p.pre.max_iters == 3
    assert yaml_params.amg_kwargs(roles)["coarse_size"] == 500
