#!/usr/bin/env python3
"""Summarize rocprofv3 CSV output per kernel.

  python profsum.py stats DIR      # from --kernel-trace --stats (kernel_stats.csv)
  python profsum.py pmc DIR        # from --pmc ... (counter_collection.csv)

`stats` prints calls, average and total duration per kernel; `pmc` prints,
per kernel and counter, the number of dispatches and the average counter
value per dispatch (FETCH_SIZE / WRITE_SIZE are in KiB per dispatch).
Kernel names are shortened to the function name and template arguments.
"""
import csv
import glob
import os
import re
import sys
from collections import defaultdict


def short(name):
    name = re.sub(r"\(.*\)$", "", name.strip())      # drop the parameter list
    name = name.replace("void ", "")
    return name[:110]


def find(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        sys.exit(f"no {pattern} under {d}")
    return hits


def stats(d):
    rows = []
    for path in find(d, "*kernel_stats.csv"):
        with open(path) as fh:
            rows += list(csv.DictReader(fh))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print(f"{'kernel':<112} {'calls':>8} {'avg_us':>10} {'total_ms':>10} {'pct':>6}")
    for r in rows:
        print(f"{short(r['Name']):<112} {int(r['Calls']):>8} {float(r['AverageNs']) / 1e3:>10.2f} "
              f"{float(r['TotalDurationNs']) / 1e6:>10.2f} {float(r['Percentage']):>6.2f}")


def pmc(d):
    acc = defaultdict(lambda: [0, 0.0])
    for path in find(d, "*counter_collection.csv"):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                key = (short(r["Kernel_Name"]), r["Counter_Name"])
                acc[key][0] += 1
                acc[key][1] += float(r["Counter_Value"])
    print(f"{'kernel':<112} {'counter':>12} {'dispatches':>10} {'avg_per_dispatch':>18}")
    for (k, c), (n, s) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        print(f"{k:<112} {c:>12} {n:>10} {s / n:>18.1f}")


if __name__ == "__main__":
    if len(sys.argv) != 3 or sys.argv[1] not in ("stats", "pmc"):
        sys.exit(__doc__)
    (stats if sys.argv[1] == "stats" else pmc)(sys.argv[2])
