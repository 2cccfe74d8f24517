"""Per-launch averages of rocprofv3 counters (`--pmc ... --output-format csv`) for one
kernel at one grid size, beside the same counters for the library copy kernel that
`solve_profile.py` launches (`k_convert<double,double>`, its largest dispatches).  Counter
values of a dispatch are summed over the rows rocprofv3 writes for it (one row per
counter instance); launches that found the solve done (under half the largest value)
are dropped.  Prints a markdown table.

usage: python pmc_counters.py <counter_collection.csv> [...] \
           --kernel 'k_spmv<double,u16,32,OP_CHEB>' --workgroups 16384"""
import argparse
import csv
from collections import defaultdict

from trace_summary import short

COPY = "k_convert<double,double>"


def dispatches(path):
    """{(dispatch id): [kernel, workgroups, {counter: value}]}"""
    out = defaultdict(lambda: [None, 0, defaultdict(float)])
    with open(path) as f:
        for r in csv.DictReader(f):
            d = out[r["Dispatch_Id"]]
            d[0] = short(r["Kernel_Name"])
            d[1] = int(r["Grid_Size"]) // max(1, int(r["Workgroup_Size"]))
            d[2][r["Counter_Name"]] += float(r["Counter_Value"])
    return out


def average(ds, pick):
    """per-counter mean over the picked dispatches (dropping near-empty ones)"""
    sel = [c for k, g, c in ds.values() if pick(k, g)]
    if not sel:
        return {}, 0
    names = sorted({n for c in sel for n in c})
    ref = names[0]
    top = max(c[ref] for c in sel)
    sel = [c for c in sel if c[ref] > 0.5 * top]
    return {n: sum(c[n] for c in sel) / len(sel) for n in names}, len(sel)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("csv", nargs="+")
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--workgroups", type=int, required=True)
    a = ap.parse_args()
    print(f"| counter | `{a.kernel}` ({a.workgroups} workgroups) | copy kernel | launches |")
    print("|---|---|---|---|")
    for path in a.csv:
        ds = dispatches(path)
        kern, nk = average(ds, lambda k, g: k == a.kernel and g == a.workgroups)
        copy_all = [c for k, g, c in ds.values() if k == COPY]
        copy, nc = ({}, 0)
        if copy_all:
            ref = sorted(copy_all[0])[0]
            top = max(c[ref] for c in copy_all)
            big = [c for c in copy_all if c[ref] > 0.9 * top]
            copy = {n: sum(c[n] for c in big) / len(big) for n in big[0]}
            nc = len(big)
        for n in sorted(kern):
            cv = copy.get(n)
            print(f"| {n} | {kern[n]:,.0f} | {'' if cv is None else f'{cv:,.0f}'} | {nk} / {nc} |")


if __name__ == "__main__":
    main()
