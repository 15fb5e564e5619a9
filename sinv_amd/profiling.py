"""Summaries of rocprofv3 output for the hot-path kernels.

    python -m sinv_amd.profiling <rocprofv3 output dir> [--range NAME]
    python -m sinv_amd.profiling --traffic FETCH_DIR WRITE_DIR --calib-bytes BYTES

The second form reads two ``rocprofv3 --pmc`` runs (``FETCH_SIZE``, then
``WRITE_SIZE``, both in KiB per dispatch) of ``bench.py --traffic-pass`` and
reports the mean HBM bytes per dispatch of each hot-path kernel, each counter
scaled by the factor that makes the calibration kernel (one elementwise
``neg`` over BYTES read and BYTES written) read exactly its known size.

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
    name = name.replace("(anonymous namespace)::", "")
    # every tile configuration of the GEMM counts as one kernel per precision
    for key, short in (("gemm_kernel<double", "gemm_kernel<double>"),
                       ("gemm_kernel<float", "gemm_kernel<float>")):
        if key in name:
            return short
    for key in ("gemm_i8x_o2_kernel", "gemm_i8x_kernel", "split_rows_kernel",
                "slice_piece_kernel", "piece_max_kernel", "inv_leaf_kernel", "gj_flow_kernel", "gj_step_kernel", "gj_panel_global_kernel",
                "gj_panel_kernel", "gj_extract_kernel", "gj_copy_kernel", "gj_permute_kernel",
                "axpby_shift_kernel", "transpose_kernel", "fill_u64_kernel", "fill_kernel",
                "ns_init_kernel", "ns_check_kernel", "ns_finish_kernel", "ns_copy_kernel",
                "dominance_kernel"):
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
    agg = {}
    for r in csv.DictReader(open(f)):
        a = agg.setdefault(_short(r["Name"]), [0, 0.0, 0.0])
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"]) / 1e6
        a[2] += float(r["Percentage"])
    return [(k, n, t, t / n * 1e3, pct) for k, (n, t, pct) in
            sorted(agg.items(), key=lambda kv: -kv[1][1])]


def _ranges(d, name):
    f = _one(os.path.join(d, "**", "*marker_api_trace.csv")) or _one(os.path.join(d, "*marker_api_trace.csv"))
    if not f:
        return []
    spans = []
    for r in csv.DictReader(open(f)):
        # exact match: "sinv_variant:rgf" must not select "sinv_variant:rgf2sided"
        if any(str(v).strip() == name for v in r.values()):
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


def pmc_means(d, counter):
    """Mean value of `counter` per dispatch, by kernel (summed over the rows of one dispatch)."""
    f = (_one(os.path.join(d, "**", "*counter_collection.csv"))
         or _one(os.path.join(d, "*counter_collection.csv")))
    if not f:
        return {}
    per = defaultdict(float)
    name = {}
    for r in csv.DictReader(open(f)):
        if r.get("Counter_Name") != counter:
            continue
        k = r.get("Dispatch_Id") or r.get("Correlation_Id")
        per[k] += float(r["Counter_Value"])
        name[k] = r["Kernel_Name"]
    agg = defaultdict(lambda: [0, 0.0])
    for k, v in per.items():
        a = agg[_short(name[k])]
        a[0] += 1
        a[1] += v
    return {k: (n, t / n) for k, (n, t) in agg.items()}


def traffic_table(fetch_dir, write_dir, calib_bytes):
    fetch = pmc_means(fetch_dir, "FETCH_SIZE")
    write = pmc_means(write_dir, "WRITE_SIZE")
    cal = [k for k in fetch if "neg" in k]
    if not cal or cal[0] not in write:
        raise SystemExit("calibration kernel (neg) not found in both runs")
    kf = calib_bytes / (fetch[cal[0]][1] * 1024.0)
    kw = calib_bytes / (write[cal[0]][1] * 1024.0)
    rows = []
    for k in sorted(fetch, key=lambda x: -fetch[x][0]):
        if k in write:
            n, f = fetch[k]
            rows.append((k, n, f * 1024.0 * kf, write[k][1] * 1024.0 * kw))
    return kf, kw, rows


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "--traffic":
        ap = argparse.ArgumentParser()
        ap.add_argument("--traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
        ap.add_argument("--calib-bytes", type=float, required=True)
        a = ap.parse_args(argv)
        kf, kw, rows = traffic_table(a.traffic[0], a.traffic[1], a.calib_bytes)
        print(f"# HBM traffic per dispatch from FETCH_SIZE ({a.traffic[0]}) and WRITE_SIZE ({a.traffic[1]})")
        print(f"calibration (neg over {a.calib_bytes:.0f} B): fetch factor {kf:.4f}, write factor {kw:.4f}")
        print(f"{'kernel':28s} {'dispatches':>10s} {'fetch_MB':>10s} {'write_MB':>10s} {'total_MB':>10s}")
        for k, n, f, w in rows:
            print(f"{k:28s} {n:10d} {f / 1e6:10.3f} {w / 1e6:10.3f} {(f + w) / 1e6:10.3f}")
        return 0
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
