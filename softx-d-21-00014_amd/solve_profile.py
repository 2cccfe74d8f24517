"""Per-kernel accounting of one benchmark solve: builds the operator and AMG hierarchy,
runs the solve once to tolerance (hipGraph replays) for its iteration count, then the same
number of iterations with every launch bracketed by HIP events (Solver.profile), and
prints, per kernel and hierarchy level, the launches, algorithmic bytes and microseconds
per launch, achieved GB/s and the fraction of the 8 TB/s HBM peak; then the measured
ceiling of the library's copy kernel over one multivector.  Launch order and kernels are
those of the timed solve (graph replay only removes launch overhead), so a
`rocprofv3 --kernel-trace` of this command gives the same kernels' durations.

--format md prints a markdown table (the profiles/ summaries), json one JSON line."""
import argparse
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pyxamg as X  # noqa: E402

ap = argparse.ArgumentParser()
ap.add_argument("--op", choices=["cube", "q1fem27", "varcoef7"], default="cube")
ap.add_argument("--nx", type=int, default=128)
ap.add_argument("--nv", type=int, default=32)
ap.add_argument("--mg-prec", type=int, default=0)
ap.add_argument("--smoother", default="Chebyshev")
ap.add_argument("--poly-order", type=int, default=4)
ap.add_argument("--format", choices=["md", "json"], default="md")
ap.add_argument("--copy-reps", type=int, default=20)
ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                help="launch-shape knob of xh_set_option, e.g. spmv_st7_enable=1")
a = ap.parse_args()
X.init(0)
for kv in a.set:
    k, v = kv.split("=")
    X.set_option(k, int(v))
N, nv = a.nx, a.nv
n = N ** 3
if a.op == "cube":
    row, col, val = X.cube(N)
    seed = 1
elif a.op == "q1fem27":
    row, col, val, _ = X.q1fem27(N, sigma=2.0, seed=1)
    seed = 5
else:
    row, col, val, _ = X.varcoef7(N, sigma=1.0, seed=2024)
    seed = 1
levels = X.Amg(row, col, val, agg_num_levels=2, num_paths=2, coarse_size=500).export()
H = X.Hier(levels)
prm = X.default_params(method="MergedPBiCGStab", precond="MultiGrid", smoother=a.smoother,
                       poly_order=a.poly_order, sweeps=1, rel_tol=1e-8, abs_tol=0.0,
                       max_iters=300, mg_prec=a.mg_prec, coarse_fp64=1 if a.mg_prec == 2 else 0,
                       graph=1)
S = X.Solver(H, prm, nv)
db = X.DeviceArray.from_host(X.rhs_uniform(seed, n, nv))
dx = X.DeviceArray(n * nv)
S.solve_device(db, dx)
iters = S.stats()[0]
dx.zero()
stats = S.profile(db, dx, iters)
rows = X.kernel_table(stats)
tot_b = sum(k["bytes"] for k in stats)
tot_ms = sum(k["ms"] for k in stats)
ms_copy = X.copy_timed(dx, db, n * nv, a.copy_reps)
copy_gbs = 16.0 * n * nv / (ms_copy * 1e-3) / 1e9
desc = (f"{a.op} {N}^3, nv={nv}, MergedPBiCGStab + AMG ({len(levels)} levels, {a.smoother}"
        f"{' order %d' % a.poly_order if a.smoother == 'Chebyshev' else ''}), mg_prec "
        f"{a.mg_prec}: init + {iters} iterations{''.join(' ' + kv for kv in a.set)}")
if a.format == "json":
    print(json.dumps(dict(workload=desc, iters=iters, kernels=rows,
                          alg_bytes_per_solve=tot_b, kernel_ms_per_solve=tot_ms,
                          copy_gb_s=round(copy_gbs, 1))))
    sys.exit(0)
print(f"# {desc}\n")
print(f"Sum over launches: {tot_b / 1e9:.3f} GB algorithmic in {tot_ms:.3f} ms of kernel time "
      f"({tot_b / (tot_ms * 1e-3) / 1e9:.1f} GB/s, {tot_b / (tot_ms * 1e-3) / X.HBM_PEAK:.3f} of "
      f"8 TB/s).  Library copy kernel over {16 * n * nv / 1e6:.0f} MB (read + write): "
      f"{copy_gbs:.1f} GB/s ({copy_gbs * 1e9 / X.HBM_PEAK:.3f} of 8 TB/s).\n")
print("| kernel | level | launches | MB/launch | µs/launch | GB/s | frac of 8 TB/s | % time |")
print("|---|---|---|---|---|---|---|---|")
for r in rows:
    lv = "Krylov" if r["level"] < 0 else str(r["level"])
    print(f"| `{r['kernel']}` | {lv} | {r['launches']} | {r['mb_per_launch']:.2f} | "
          f"{r['us_per_launch']:.1f} | {r['gb_s']:.0f} | {r['frac']:.3f} | "
          f"{100 * r['ms_total'] / tot_ms:.1f} |")
