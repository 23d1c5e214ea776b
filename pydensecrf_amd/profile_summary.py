"""Summarise rocprofv3 output (kernel-trace stats and PMC counters) as text.

    python -m pydensecrf_amd.profile_summary TRACE_DIR [--pmc FETCH_DIR WRITE_DIR SQ_DIR]

TRACE_DIR holds the *_kernel_stats.csv of `rocprofv3 --kernel-trace --stats
--output-format csv`; the optional directories hold the *_counter_collection.csv of
`rocprofv3 --pmc ...` passes.  HBM traffic per dispatch is FETCH_SIZE / WRITE_SIZE
(KiB) scaled by what the same counters read for the largest plain device copy kernel
in the run (a copy of known size: bytes moved / counter KiB).
"""
import csv
import glob
import os
import sys
from collections import defaultdict


def _short(name):
    name = name.replace("dcrf::(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0][:60]


def kernel_stats(trace_dir):
    path = sorted(glob.glob(os.path.join(trace_dir, "**", "*kernel_stats.csv"), recursive=True))
    if not path:
        raise SystemExit("no *kernel_stats.csv under %s" % trace_dir)
    rows = list(csv.DictReader(open(path[0])))
    return path[0], rows


def counters(pmc_dir):
    path = sorted(glob.glob(os.path.join(pmc_dir, "**", "*counter_collection.csv"), recursive=True))
    if not path:
        return None, {}
    acc = defaultdict(lambda: defaultdict(list))
    for r in csv.DictReader(open(path[0])):
        acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return path[0], acc


def main(argv):
    trace_dir = argv[0]
    src, rows = kernel_stats(trace_dir)
    print("kernel-trace stats: %s" % src)
    print("%-60s %8s %12s %12s %8s" % ("kernel", "calls", "avg_us", "total_ms", "pct"))
    for r in rows:
        print("%-60s %8s %12.2f %12.3f %8.2f" % (_short(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e3,
                                                 float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    if "--pmc" in argv:
        dirs = argv[argv.index("--pmc") + 1:]
        merged = defaultdict(dict)
        for d in dirs:
            s, acc = counters(d)
            print("counters: %s" % s)
            for k, cs in acc.items():
                for c, vals in cs.items():
                    merged[k][c] = sum(vals) / len(vals)
        print("%-60s %14s %14s %14s" % ("kernel (avg per dispatch)", "FETCH_KiB", "WRITE_KiB", "VALU/wave"))
        for k, cs in sorted(merged.items(), key=lambda kv: -kv[1].get("FETCH_SIZE", 0)):
            valu = cs.get("SQ_INSTS_VALU", 0) / cs["SQ_WAVES"] if cs.get("SQ_WAVES") else float("nan")
            print("%-60s %14.1f %14.1f %14.1f" % (_short(k), cs.get("FETCH_SIZE", float("nan")),
                                                  cs.get("WRITE_SIZE", float("nan")), valu))


if __name__ == "__main__":
    main(sys.argv[1:])
