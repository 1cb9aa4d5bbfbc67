#!/usr/bin/env python3
"""Summarize rocprofv3 CSV output per kernel.

  python profsum.py stats DIR      # from --kernel-trace --stats (kernel_stats.csv)
  python profsum.py pmc DIR        # from --pmc ... (counter_collection.csv)
  python profsum.py traffic RD WR COPY_BYTES
                                   # L2-to-memory bytes per dispatch from the
                                   # request counters (RD: a --pmc folder or
                                   # `pmc` summary of TCC_EA0_RDREQ_{32B,64B,
                                   # 128B}_sum and TCC_EA0_RDREQ_sum; WR: of
                                   # TCC_EA0_WRREQ_64B_sum and TCC_EA0_WRREQ_sum);
                                   # the copy kernel, which moves COPY_BYTES
                                   # each way, is printed as the check

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


def traffic(rd_dir, wr_dir, copy_bytes, copy_kernel="ifos::k_copy"):
    """bytes each kernel moves between L2 and memory per dispatch, from the
    request counters: reads 32 B x RDREQ_32B + 64 B x RDREQ_64B + 128 B x
    RDREQ_128B, writes 64 B x WRREQ_64B + 32 B x the other write requests.
    (FETCH_SIZE, as rocprofv3 derives it on this part, counts a 128-byte read
    as 64 bytes.)  The copy kernel must come out at copy_bytes each way."""
    rd, wr = _pmc_table(rd_dir), _pmc_table(wr_dir)
    avg = lambda t, k, c: (t[(k, c)][1] / t[(k, c)][0]) if (k, c) in t and t[(k, c)][0] else 0.0
    kernels = sorted({k for (k, c) in rd if c == "TCC_EA0_RDREQ_sum"})
    rows = []
    for k in kernels:
        r32, r64, r128 = (avg(rd, k, f"TCC_EA0_RDREQ_{b}_sum") for b in ("32B", "64B", "128B"))
        rall = avg(rd, k, "TCC_EA0_RDREQ_sum")
        w64, wall = avg(wr, k, "TCC_EA0_WRREQ_64B_sum"), avg(wr, k, "TCC_EA0_WRREQ_sum")
        rb = 32.0 * r32 + 64.0 * r64 + 128.0 * r128
        wb = 64.0 * w64 + 32.0 * (wall - w64)
        other = rall - r32 - r64 - r128
        rows.append((k, rb, wb, other))
    cr = [r for r in rows if r[0] == copy_kernel]
    if cr:
        print(f"# check: {copy_kernel} moves {copy_bytes} B each way per dispatch; the counters give "
              f"{cr[0][1] / copy_bytes:.4f} x read, {cr[0][2] / copy_bytes:.4f} x write")
    print(f"# read requests of no counted size (RDREQ - 32B - 64B - 128B), largest: "
          f"{max((r[3] for r in rows), default=0.0):.0f} per dispatch")
    print(f"{'kernel':<60} {'read_bytes':>16} {'write_bytes':>16} {'hbm_bytes':>16}")
    for k, rb, wb, _ in sorted(rows, key=lambda r: -(r[1] + r[2])):
        print(f"{k:<60} {rb:>16.0f} {wb:>16.0f} {rb + wb:>16.0f}")


def read_traffic(path):
    """{kernel: HBM bytes per dispatch} from a traffic summary"""
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
