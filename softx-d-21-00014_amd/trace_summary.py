"""Per-(kernel, grid size) summary of a rocprofv3 kernel trace (`--kernel-trace
--output-format csv` writes <name>_kernel_trace.csv): dispatches, average and total
microseconds, sorted by total time.  The grid size tells the hierarchy levels of one
kernel apart (the finest level has the largest grid), so the average of one level's
launches can be set against the event-timed figure of `solve_profile.py` for the same
kernel and level.

Launches enqueued behind the last iteration of a solve find it done and return at once
(a few microseconds); --min-us drops dispatches shorter than that from the averages.

usage: python trace_summary.py <kernel_trace.csv> [--top N] [--match SUBSTRING]
                               [--min-us US]"""
import argparse
import csv
import re
from collections import defaultdict

OPS = {"k_spmv": ["OP_SET", "OP_ADD", "OP_RESID", "OP_DOT1", "OP_DOT3", "OP_CHEB0", "OP_CHEB",
                  "OP_PROLONG32", "OP_OFFBLOCK", "OP_RESID0", "OP_RDOT4", "OP_CHEBZ",
                  "OP_CHEBR"]}
VEC = ["V_S_UPDATE", "V_XR_UPDATE", "V_P_UPDATE", "V_GROUP1", "V_X_UPDATE", "V_R_UPDATE",
       "V_GROUP5", "V_DOT3", "V_DOT1", "V_CG_XR", "V_DOT2", "V_CG_P", "V_COPY_DOT", "V_COPY",
       "V_RB_UPDATE", "V_RB_UPDATE_MERGED"]


def short(name):
    """kernel template instance with the op enums named, without namespace and arguments"""
    m = re.search(r"(k_\w+)<([^()]*)>", name)
    if not m:
        return name.split("(")[0]
    k, args = m.group(1), [a.strip() for a in m.group(2).split(",")]
    args = ["u16" if a == "unsigned short" else "u32" if a == "unsigned int" else a
            for a in args]
    if k == "k_spmv" and len(args) == 5 and args[3].isdigit():
        # k_spmv<VT, IT, NV, OP, BX>: the x loads per batch shown only when not 8, as the
        # solver's launch accounting names the kernel
        args = args[:3] + [OPS["k_spmv"][int(args[3])]] + ([args[4]] if args[4] != "8" else [])
    elif k.startswith("k_spmv") and args[-1].isdigit():
        args[-1] = OPS["k_spmv"][int(args[-1])]
    if k == "k_vec" and args[-1].isdigit():
        args[-1] = VEC[int(args[-1])]
    return f"{k}<{','.join(args)}>"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--match", default="")
    ap.add_argument("--min-us", type=float, default=0.0)
    a = ap.parse_args()
    acc = defaultdict(lambda: [0, 0.0])
    with open(a.trace) as f:
        for r in csv.DictReader(f):
            name = short(r["Kernel_Name"])
            if a.match and a.match not in name:
                continue
            grid = int(r["Grid_Size_X"]) // max(1, int(r["Workgroup_Size_X"]))
            us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            if us < a.min_us:
                continue
            acc[(name, grid)][0] += 1
            acc[(name, grid)][1] += us
    rows = sorted(acc.items(), key=lambda kv: -kv[1][1])[: a.top]
    print("| kernel | workgroups | dispatches | avg µs | total ms |")
    print("|---|---|---|---|---|")
    for (name, grid), (n, us) in rows:
        print(f"| `{name}` | {grid} | {n} | {us / n:.1f} | {us / 1e3:.2f} |")


if __name__ == "__main__":
    main()
