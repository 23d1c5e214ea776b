"""Summarise rocprofv3 output (kernel-trace stats and PMC counters) as text.

    python -m pydensecrf_amd.profile_summary TRACE_DIR [--copy-kib K] [--pmc FETCH_DIR WRITE_DIR SQ_DIR]

TRACE_DIR holds the *_kernel_stats.csv of `rocprofv3 --kernel-trace --stats
--output-format csv`; the optional directories hold the *_counter_collection.csv of
`rocprofv3 --pmc ...` passes.  HBM traffic per dispatch is FETCH_SIZE / WRITE_SIZE
(KiB) scaled by what the same counters read for the largest plain device copy kernel
in the run, a copy of K KiB (bench.py --copy-calibration copies 262144 KiB three
times): scale = K / counter reading.  With --copy-kib the summary ends with lines
    traffic <class> <HBM bytes per launch>
for the mean-field kernel classes (of the kernels whose name starts with the class's
kernel prefix, the one with the most dispatches).
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


CLASSES = {"splat": "k_splat", "combine": "k_combine", "blur": "k_blur", "meanfield": "k_meanfield"}
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def counters(pmc_dir):
    path = sorted(glob.glob(os.path.join(pmc_dir, "**", "*counter_collection.csv"), recursive=True))
    if not path:
        return None, {}, {}
    acc = defaultdict(lambda: defaultdict(list))
    copy_max = {}
    for r in csv.DictReader(open(path[0])):
        c, v = r["Counter_Name"], float(r["Counter_Value"])
        acc[r["Kernel_Name"]][c].append(v)
        if r["Kernel_Name"] == COPY_KERNEL:
            copy_max[c] = max(copy_max.get(c, 0.0), v)
    return path[0], acc, copy_max


def main(argv):
    trace_dir = argv[0]
    src, rows = kernel_stats(trace_dir)
    print("kernel-trace stats: %s" % src)
    print("%-60s %8s %12s %12s %8s" % ("kernel", "calls", "avg_us", "total_ms", "pct"))
    for r in rows:
        print("%-60s %8s %12.2f %12.3f %8.2f" % (_short(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e3,
                                                 float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    copy_kib = float(argv[argv.index("--copy-kib") + 1]) if "--copy-kib" in argv else None
    if "--pmc" in argv:
        dirs = argv[argv.index("--pmc") + 1:]
        merged = defaultdict(dict)
        calls = defaultdict(int)
        copy_max = {}
        for d in dirs:
            s, acc, cm = counters(d)
            print("counters: %s" % s)
            copy_max.update(cm)
            for k, cs in acc.items():
                for c, vals in cs.items():
                    merged[k][c] = sum(vals) / len(vals)
                    calls[k] = max(calls[k], len(vals))
        print("%-60s %14s %14s %14s" % ("kernel (avg per dispatch)", "FETCH_KiB", "WRITE_KiB", "VALU/wave"))
        for k, cs in sorted(merged.items(), key=lambda kv: -kv[1].get("FETCH_SIZE", 0)):
            valu = cs.get("SQ_INSTS_VALU", 0) / cs["SQ_WAVES"] if cs.get("SQ_WAVES") else float("nan")
            print("%-60s %14.1f %14.1f %14.1f" % (_short(k), cs.get("FETCH_SIZE", float("nan")),
                                                  cs.get("WRITE_SIZE", float("nan")), valu))
        # every other counter of the passes, per dispatch, for each mean-field kernel class
        # (the kernel of the class with the most dispatches)
        shown = {"FETCH_SIZE", "WRITE_SIZE", "SQ_WAVES", "SQ_INSTS_VALU"}
        others = sorted({c for cs in merged.values() for c in cs} - shown)
        if others:
            print("%-40s %s" % ("counter (avg per dispatch)", " ".join("%16s" % c for c in CLASSES)))
            rep = {}
            for cls, prefix in CLASSES.items():
                names = [k for k in merged if _short(k).startswith(prefix)]
                rep[cls] = merged[max(names, key=lambda n: calls[n])] if names else {}
            for c in others:
                print("%-40s %s" % (c, " ".join("%16.4g" % rep[cls].get(c, float("nan")) for cls in CLASSES)))
        if copy_kib and copy_max.get("FETCH_SIZE") and copy_max.get("WRITE_SIZE"):
            fs, ws = copy_kib / copy_max["FETCH_SIZE"], copy_kib / copy_max["WRITE_SIZE"]
            print("copy calibration: %.0f KiB copy reads FETCH_SIZE %.1f, WRITE_SIZE %.1f -> scales %.3f, %.3f"
                  % (copy_kib, copy_max["FETCH_SIZE"], copy_max["WRITE_SIZE"], fs, ws))
            for cls, prefix in CLASSES.items():
                names = [k for k in merged if _short(k).startswith(prefix)]
                if not names:
                    continue
                k = max(names, key=lambda n: calls[n])
                cs = merged[k]
                if "FETCH_SIZE" in cs and "WRITE_SIZE" in cs:
                    b = (cs["FETCH_SIZE"] * fs + cs["WRITE_SIZE"] * ws) * 1024
                    print("traffic %s %d" % (cls, round(b)))


if __name__ == "__main__":
    main(sys.argv[1:])
