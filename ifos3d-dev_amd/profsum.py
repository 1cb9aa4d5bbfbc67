#!/usr/bin/env python3
"""Summarize rocprofv3 CSV output per kernel.

  python profsum.py stats DIR      # from --kernel-trace --stats (kernel_stats.csv)
  python profsum.py pmc DIR        # from --pmc ... (counter_collection.csv)
  python profsum.py traffic FETCH WRITE COPY_BYTES
                                   # HBM bytes per dispatch, calibrated by the
                                   # copy kernel (FETCH/WRITE: a --pmc folder
                                   # or a `pmc` summary)

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


def _pmc_table(d):
    """{(kernel, counter): [dispatches, sum]} from a --pmc output folder or
    from a summary printed by `pmc`"""
    acc = defaultdict(lambda: [0, 0.0])
    if os.path.isfile(d):
        with open(d) as fh:
            for line in fh:
                if line.startswith("#"):
                    continue
                parts = line.rsplit(None, 3)
                if len(parts) == 4 and parts[2].isdigit():
                    n, avg = int(parts[2]), float(parts[3])
                    acc[(parts[0].strip(), parts[1])] = [n, n * avg]
        return acc
    for path in find(d, "*counter_collection.csv"):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                key = (short(r["Kernel_Name"]), r["Counter_Name"])
                acc[key][0] += 1
                acc[key][1] += float(r["Counter_Value"])
    return acc


def pmc(d):
    acc = _pmc_table(d)
    print(f"{'kernel':<112} {'counter':>12} {'dispatches':>10} {'avg_per_dispatch':>18}")
    for (k, c), (n, s) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        print(f"{k:<112} {c:>12} {n:>10} {s / n:>18.1f}")


def traffic(fetch_dir, write_dir, copy_bytes, copy_kernel="ifos::k_copy"):
    """FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of each kernel, each
    scaled by copy_bytes / (what the counter reads for the copy kernel,
    which reads and writes copy_bytes per dispatch)"""
    f = {k: n and s / n for (k, c), (n, s) in _pmc_table(fetch_dir).items() if c == "FETCH_SIZE"}
    w = {k: n and s / n for (k, c), (n, s) in _pmc_table(write_dir).items() if c == "WRITE_SIZE"}
    fs = copy_bytes / (f[copy_kernel] * 1024.0)
    ws = copy_bytes / (w[copy_kernel] * 1024.0)
    print(f"# calibration: {copy_kernel} moves {copy_bytes} B each way per dispatch; "
          f"FETCH_SIZE scale {fs:.4f}, WRITE_SIZE scale {ws:.4f}")
    print(f"{'kernel':<60} {'fetch_KiB':>14} {'write_KiB':>14} {'hbm_bytes':>16}")
    for k in sorted(f, key=lambda k: -(f[k] * fs + w.get(k, 0.0) * ws)):
        hb = (f[k] * fs + w.get(k, 0.0) * ws) * 1024.0
        print(f"{k:<60} {f[k]:>14.1f} {w.get(k, 0.0):>14.1f} {hb:>16.0f}")


def read_traffic(path):
    """{kernel: calibrated HBM bytes per dispatch} from a traffic summary"""
    out = {}
    with open(path) as fh:
        for line in fh:
            if line.startswith("#") or line.startswith("kernel") or not line.strip():
                continue
            parts = line.rsplit(None, 3)
            if len(parts) == 4:
                try:
                    out[parts[0].strip()] = float(parts[3])
                except ValueError:
                    pass
    return out


if __name__ == "__main__":
    if len(sys.argv) == 5 and sys.argv[1] == "traffic":
        traffic(sys.argv[2], sys.argv[3], int(sys.argv[4]))
    elif len(sys.argv) == 3 and sys.argv[1] in ("stats", "pmc"):
        (stats if sys.argv[1] == "stats" else pmc)(sys.argv[2])
    else:
        sys.exit(__doc__)
