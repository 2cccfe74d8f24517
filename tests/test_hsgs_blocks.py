"""Hybrid GS block count against the reference's per-core blocks (CPU, oracle only).

The reference's HSGS is sequential inside each core's block of rows and Jacobi between
blocks (src/blas2/blas2.h:185-281, block sizes from split_range, src/misc/misc.h:46-55):
its iteration count depends on the core count it runs on.  The GPU uses blocks of
split_range(n, ceil(n/64)) rows.  On the 27-point operator with lognormal k (sigma = 2),
MergedPBiCGStab + AMG with HSGS 1+1 on every level, the oracle runs the same solve with
one block (exact symmetric GS, the reference on one core), 16 and 128 core blocks, and the
GPU's 64-row blocks; the GPU's block count must converge no slower than the reference on a
16- or 128-core node (within one iteration), and exact GS no slower than any of them."""
import numpy as np
import pytest


@pytest.mark.parametrize("N", [24, 48])
def test_hsgs_block_count_against_core_blocks(X, O, N):
    row, col, val, _ = X.q1fem27(N, sigma=2.0, seed=1)
    levels = X.Amg(row, col, val, agg_num_levels=2, num_paths=2, coarse_size=500).export()
    b = X.rhs_uniform(5, N ** 3, 2)
    H = O.Hier(levels)
    its = {}
    for label, kw in (("1 core", dict(hsgs_block_rows=0, hsgs_ncores=1)),
                      ("16 cores", dict(hsgs_block_rows=0, hsgs_ncores=16)),
                      ("128 cores", dict(hsgs_block_rows=0, hsgs_ncores=128)),
                      ("64-row blocks", dict(hsgs_block_rows=64))):
        prm = X.default_params(method="MergedPBiCGStab", smoother="HSGS", sweeps=1,
                               max_iters=300, rel_tol=1e-8)
        for k, v in kw.items():
            setattr(prm, k, v)
        r = O.solve(H, O.params_from(prm), 2, b)
        assert r["conv"].all(), label
        its[label] = r["iters"]
    print(f"q1fem27 {N}^3 sigma 2, HSGS iterations: {its}")
    assert its["64-row blocks"] <= min(its["16 cores"], its["128 cores"]) + 1
    assert its["1 core"] <= min(its.values())
