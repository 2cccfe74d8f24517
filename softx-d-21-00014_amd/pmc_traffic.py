"""HBM traffic per launch of one kernel from two rocprofv3 counter runs of
`solve_profile.py` (`--pmc FETCH_SIZE`, then `--pmc WRITE_SIZE`, csv output), scaled by
what the same counters read for the library copy kernel of known traffic that the same
command launches (`xh_copy_device_timed` over n*nv doubles: 8*n*nv bytes read and
written).  Prints one JSON object (kernel, level, bytes, calibration, the git object id of
xh_spmv.hip the measurement belongs to) for profiles/.

usage: python pmc_traffic.py <fetch counter_collection.csv> <write counter_collection.csv>
           --kernel 'k_spmv<double,u16,32,OP_CHEB>' --workgroups 16384 --copy-bytes B
           --alg-bytes A --level 0"""
import argparse
import csv
import hashlib
import json
import os
from collections import defaultdict

from trace_summary import short


def per_dispatch(path, counter):
    """{dispatch id: (kernel, workgroups, value in KiB)} of one counter"""
    out = defaultdict(lambda: [None, 0, 0.0])
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            d = out[r["Dispatch_Id"]]
            d[0] = short(r["Kernel_Name"])
            d[1] = int(r["Grid_Size"]) // max(1, int(r["Workgroup_Size"]))
            d[2] += float(r["Counter_Value"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch")
    ap.add_argument("write")
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--workgroups", type=int, required=True)
    ap.add_argument("--level", type=int, default=0)
    ap.add_argument("--copy-bytes", type=float, required=True,
                    help="bytes the copy kernel reads (= bytes it writes)")
    ap.add_argument("--alg-bytes", type=float, required=True)
    a = ap.parse_args()
    res = {}
    for path, counter in ((a.fetch, "FETCH_SIZE"), (a.write, "WRITE_SIZE")):
        d = per_dispatch(path, counter)
        copy = [v for k, g, v in d.values() if k == "k_convert<double,double>"]
        copy = max(copy)  # the multivector copies (the largest dispatches of that kernel)
        kern = [v for k, g, v in d.values() if k == a.kernel and g == a.workgroups]
        top = max(kern)
        kern = [v for v in kern if v > 0.5 * top]  # drop launches that found the solve done
        res[counter] = (sum(kern) / len(kern), len(kern), copy)
    f_kib, nf, f_copy = res["FETCH_SIZE"]
    w_kib, nw, w_copy = res["WRITE_SIZE"]
    fs = a.copy_bytes / (f_copy * 1024.0)
    ws = a.copy_bytes / (w_copy * 1024.0)
    hbm = f_kib * 1024.0 * fs + w_kib * 1024.0 * ws
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "hip", "xh_spmv.hip")
    with open(src, "rb") as f:
        data = f.read()
    print(json.dumps({
        "kernel": a.kernel, "level": a.level, "workgroups": a.workgroups,
        "fetch_size_kib_per_launch": round(f_kib, 1), "write_size_kib_per_launch": round(w_kib, 1),
        "launches_averaged": [nf, nw],
        "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
        "calibration": f"library copy kernel over {a.copy_bytes:.0f} bytes read + written: "
                       f"FETCH_SIZE {f_copy:.1f} KiB, WRITE_SIZE {w_copy:.1f} KiB per dispatch",
        "hbm_bytes_per_launch": int(hbm),
        "algorithmic_bytes_per_launch": int(a.alg_bytes),
        "traffic_over_algorithmic": round(hbm / a.alg_bytes, 4),
        "xh_spmv_blob": hashlib.sha1(b"blob %d\0" % len(data) + data).hexdigest(),
    }, indent=1))


if __name__ == "__main__":
    main()
