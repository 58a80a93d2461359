"""Summaries of rocprofv3 output for the library's kernels.

    rocprofv3 --kernel-trace --stats --output-format csv -d DIR -o run -- \
        python bench.py --steps 2 --warmup 1 --calls-log DIR/calls.json
    python -m spaghetty_amd.profiling DIR/run_kernel_trace.csv \
        --stats DIR/run_kernel_stats.csv --calls DIR/calls.json

prints the per-kernel statistics of the library's kernels and, with the
bench's launch list, the average duration of the dominant kernel's launches
(the 100^4 sorts of the timed steps) so it can be set beside the figure
bench.py measured with HIP events.
"""
from __future__ import annotations

import argparse
import csv
import json
import sys

OURS = ("k_vtile", "k_tiled", "k_rows", "k_batch")


def _col(row, *names):
    for n in names:
        if n in row:
            return row[n]
    raise KeyError(names)


def load_trace(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            name = _col(r, "Kernel_Name", "KernelName", "Name")
            t0 = int(_col(r, "Start_Timestamp", "BeginNs", "Start"))
            t1 = int(_col(r, "End_Timestamp", "EndNs", "End"))
            did = int(_col(r, "Dispatch_Id", "DispatchId", "Index"))
            rows.append({"name": name, "dur_ns": t1 - t0, "id": did, "start": t0})
    rows.sort(key=lambda r: r["id"])
    return rows


def short(name: str) -> str:
    for k in OURS:
        if k in name:
            i = name.find(k)
            return name[i:].split("(")[0]
    return name[:60]


def summarize(trace, stats=None, calls=None, out=sys.stdout):
    ours = [r for r in trace if any(k in r["name"] for k in OURS)]
    agg = {}
    for r in ours:
        a = agg.setdefault(short(r["name"]), [0, 0])
        a[0] += 1
        a[1] += r["dur_ns"]
    print("kernel dispatches of the library (from the kernel trace):", file=out)
    print(f"{'kernel':40s} {'calls':>7s} {'total_us':>12s} {'avg_us':>10s}", file=out)
    for k, (n, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"{k:40s} {n:7d} {t / 1e3:12.1f} {t / n / 1e3:10.2f}", file=out)
    if stats:
        print("\nrocprofv3 --stats (all kernels, top 12 by total time):", file=out)
        with open(stats) as f:
            rs = list(csv.DictReader(f))
        rs.sort(key=lambda r: -float(_col(r, "TotalDurationNs", "Total_Duration_Ns")))
        for r in rs[:12]:
            print(f"  {short(_col(r, 'Name', 'KernelName')):40s} calls={_col(r, 'Calls')} "
                  f"avg_us={float(_col(r, 'AverageNs', 'Average_Ns')) / 1e3:.2f} "
                  f"total_us={float(_col(r, 'TotalDurationNs', 'Total_Duration_Ns')) / 1e3:.1f}",
                  file=out)
    if calls:
        with open(calls) as f:
            c = json.load(f)
        seq = c["calls"]
        per_step = len(seq)
        sweep = [r for r in ours if "k_batch" not in r["name"]]
        first = c["warmup"] * per_step
        timed = sweep[first:first + c["steps"] * per_step]
        dom = []
        for i, r in enumerate(timed):
            dims, perm, f_, beta = seq[i % per_step]
            if dims == [100, 100, 100, 100] and beta == 0.0:
                dom.append(r["dur_ns"])
        if dom:
            avg = sum(dom) / len(dom)
            gbs = 16 * 100 ** 4 / (avg * 1e-9) / 1e9
            print(f"\ndominant kernel (tce_sort_4 100^4 sorts of the timed steps): "
                  f"{len(dom)} dispatches, avg {avg / 1e3:.2f} us -> {gbs:.1f} GB/s "
                  f"algorithmic ({100 * gbs / 8000:.1f}% of 8 TB/s)", file=out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--stats")
    ap.add_argument("--calls")
    a = ap.parse_args()
    summarize(load_trace(a.trace), a.stats, a.calls)


if __name__ == "__main__":
    main()
