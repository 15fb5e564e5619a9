"""Summaries of rocprofv3 output for the hot-path kernels.

    python -m sinv_amd.profiling <rocprofv3 output dir> [--range NAME]

Reads the ``*_kernel_stats.csv`` (per-kernel totals over the whole run) and,
when present, ``*_kernel_trace.csv`` plus ``*_marker_api_trace.csv`` to give
the average dispatch duration of each kernel inside a roctx range (bench.py
marks its timed region ``sinv_bench_timed:<mode>``).  The output is a short
text table meant to be kept next to the benchmark figures it backs.
"""

import argparse
import csv
import glob
import os
import sys
from collections import defaultdict


def _short(name):
    for key in ("gemm_kernel<double>", "gemm_kernel<float>", "inv_leaf_kernel", "axpby_shift_kernel",
                "transpose_kernel", "fill_kernel", "fill_u64_kernel"):
        if key in name:
            return key
    return name.split("(")[0][-60:]


def _one(pattern):
    hits = sorted(glob.glob(pattern))
    return hits[0] if hits else None


def stats_table(d):
    f = _one(os.path.join(d, "**", "*kernel_stats.csv")) or _one(os.path.join(d, "*kernel_stats.csv"))
    if not f:
        return []
    rows = list(csv.DictReader(open(f)))
    out = []
    for r in rows:
        out.append((_short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]) / 1e6,
                    float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    return out


def _ranges(d, name):
    f = _one(os.path.join(d, "**", "*marker_api_trace.csv")) or _one(os.path.join(d, "*marker_api_trace.csv"))
    if not f:
        return []
    spans = []
    for r in csv.DictReader(open(f)):
        txt = " ".join(str(v) for v in r.values())
        if name in txt:
            try:
                spans.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
            except (KeyError, ValueError):
                pass
    return spans


def range_table(d, name):
    spans = _ranges(d, name)
    f = _one(os.path.join(d, "**", "*kernel_trace.csv")) or _one(os.path.join(d, "*kernel_trace.csv"))
    if not spans or not f:
        return []
    lo = min(s for s, _ in spans)
    hi = max(e for _, e in spans)
    agg = defaultdict(lambda: [0, 0.0])
    for r in csv.DictReader(open(f)):
        s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if s >= lo and e <= hi:
            a = agg[_short(r["Kernel_Name"])]
            a[0] += 1
            a[1] += (e - s) / 1e3
    return [(k, n, t / 1e3, t / n) for k, (n, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--range", action="append", default=[])
    a = ap.parse_args(argv)
    print(f"# rocprofv3 summary of {a.dir}")
    print("## all dispatches (kernel_stats)")
    print(f"{'kernel':28s} {'calls':>8s} {'total_ms':>10s} {'avg_us':>10s} {'pct':>7s}")
    for k, n, t, avg, pct in stats_table(a.dir):
        print(f"{k:28s} {n:8d} {t:10.2f} {avg:10.2f} {pct:7.2f}")
    for name in a.range:
        rows = range_table(a.dir, name)
        print(f"## dispatches inside roctx range '{name}'")
        if not rows:
            print("(range or trace not found)")
            continue
        print(f"{'kernel':28s} {'calls':>8s} {'total_ms':>10s} {'avg_us':>10s}")
        for k, n, t, avg in rows:
            print(f"{k:28s} {n:8d} {t:10.2f} {avg:10.2f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
