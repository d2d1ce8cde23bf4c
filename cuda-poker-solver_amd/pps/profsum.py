"""Short text summaries of rocprofv3 CSV output (--output-format csv).

    python -m pps.profsum stats DIR      kernel stats: calls, total and mean
                                         duration per kernel
    python -m pps.profsum counters DIR   counter collection: mean value per
                                         dispatch for each (kernel, counter)

DIR is the -d directory of the rocprofv3 run; every matching CSV under it
is read (*_kernel_stats.csv, *_counter_collection.csv).
"""
import collections
import csv
import glob
import os
import sys


def _files(root, suffix):
    return sorted(glob.glob(os.path.join(root, "**", f"*{suffix}"), recursive=True))


def _short(name, width=64):
    """Kernel name without return type, anonymous namespace and arguments."""
    name = name.replace("(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    name = name.split("(")[0]
    return name if len(name) <= width else name[:width - 3] + "..."


def stats(root):
    rows = []
    for path in _files(root, "_kernel_stats.csv"):
        with open(path) as f:
            rows += list(csv.DictReader(f))
    if not rows:
        raise SystemExit(f"no *_kernel_stats.csv under {root}")
    agg = collections.defaultdict(lambda: [0, 0.0])
    for r in rows:
        a = agg[r["Name"]]
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"])
    tot = sum(a[1] for a in agg.values())
    print(f"{'kernel':64s} {'calls':>8s} {'total_ms':>10s} {'mean_us':>9s} {'pct':>6s}")
    for name, (calls, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"{_short(name):64s} {calls:8d} {ns / 1e6:10.3f} {ns / calls / 1e3:9.2f} "
              f"{100 * ns / tot:6.1f}")


def counters(root):
    rows = []
    for path in _files(root, "_counter_collection.csv"):
        with open(path) as f:
            rows += list(csv.DictReader(f))
    if not rows:
        raise SystemExit(f"no *_counter_collection.csv under {root}")
    per = collections.defaultdict(float)  # (kernel, counter, dispatch) -> value
    for r in rows:
        per[(r["Kernel_Name"], r["Counter_Name"], r["Dispatch_Id"])] += float(r["Counter_Value"])
    agg = collections.defaultdict(lambda: [0, 0.0])
    for (k, c, _), v in per.items():
        a = agg[(k, c)]
        a[0] += 1
        a[1] += v
    print(f"{'kernel':64s} {'counter':12s} {'dispatches':>10s} {'mean/dispatch':>14s}")
    for (k, c), (n, v) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"{_short(k):64s} {c:12s} {n:10d} {v / n:14.1f}")


if __name__ == "__main__":
    if len(sys.argv) != 3 or sys.argv[1] not in ("stats", "counters"):
        raise SystemExit(__doc__)
    (stats if sys.argv[1] == "stats" else counters)(sys.argv[2])
