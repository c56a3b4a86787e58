"""Summaries of rocprofv3 result databases (the rocpd SQLite files rocprofv3 writes).

    python -m ptrans.profiling kernels  <run_results.db>
    python -m ptrans.profiling counters <run_results.db> [kernel-substring]

'kernels' prints per-kernel call count and average/total duration; 'counters' prints,
per kernel and counter, the mean value per dispatch.
"""
from __future__ import annotations

import sqlite3
import sys
from collections import defaultdict


def kernel_stats(db: str):
    """Per kernel: calls, average and total duration (ns, from the dispatch records)."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(end - start), avg(end - start) from kernels "
                       "group by name order by sum(end - start) desc").fetchall()
    tot = sum(r[2] for r in rows) or 1
    return [dict(name=r[0], calls=r[1], total_ns=r[2], average_ns=r[3], percent=100.0 * r[2] / tot)
            for r in rows]


def counter_means(db: str, kernel_filter: str = ""):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, dispatch_id, counter_name, value, "
                       "start, end from counters_collection").fetchall()
    acc = defaultdict(list)
    for name, disp, cname, val, t0, t1 in rows:
        if kernel_filter in name:
            acc[(name, cname)].append(val)
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def main(argv):
    if len(argv) < 3 or argv[1] not in ("kernels", "counters"):
        print(__doc__)
        return 2
    if argv[1] == "kernels":
        print("%-90s %6s %14s %14s %7s" % ("kernel", "calls", "average_us", "total_us", "%"))
        for k in kernel_stats(argv[2]):
            print("%-90s %6d %14.3f %14.3f %7.2f" % (k["name"][:90], k["calls"], k["average_ns"] / 1e3,
                                                   k["total_ns"] / 1e3, k["percent"]))
    else:
        filt = argv[3] if len(argv) > 3 else ""
        print("%-70s %-22s %18s %5s" % ("kernel", "counter", "mean/dispatch", "n"))
        for (name, cname), (mean, n) in sorted(counter_means(argv[2], filt).items()):
            print("%-70s %-22s %18.6g %5d" % (name[:70], cname, mean, n))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
