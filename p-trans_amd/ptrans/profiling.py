"""Summaries of rocprofv3 result databases (the rocpd SQLite files rocprofv3 writes).

    python -m ptrans.profiling kernels  <run_results.db>
    python -m ptrans.profiling counters <run_results.db> [kernel-substring]
    python -m ptrans.profiling traffic  <kt.db> <fetch.db> <write.db> <copy_fetch.db>
                                        <copy_write.db> <sq.db> <copy_mib> <paths> <workload>
                                        [key=kernel-substring[+kernel-substring] ...] [tcc=<tcc.db>]

'kernels' prints per-kernel call count and average/total duration; 'counters' prints,
per kernel and counter, the mean value per dispatch.  'traffic' prints the JSON summary
the bench's roofline reads: per trace-kernel precision its average duration (kernel
trace; a key naming "a+b" is a launch of several kernels, some of them concurrent: its
duration is the span from the first start to the last end, launches delimited by the
dispatches of kernel a, and its counters are summed over every dispatch of the launch),
HBM bytes per launch (FETCH_SIZE
and WRITE_SIZE in KiB, each scaled by what the
same counter reads for a device copy of copy_mib MiB) and VALU wave-instructions per
launch and per path with the issue fraction over 256 CUs x 2.4 GHz; with a TCC_HIT_sum /
TCC_MISS_sum run (tcc=...) also the L2 hit rate.
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


def launch_spans(db: str, subs, lead_ns: float = 20.0e6):
    """Average span (ns) of the launches made of the kernels whose names contain one of
    subs, and the number of launches.  Each dispatch of the first kernel marks a launch,
    which holds the dispatches that start from lead_ns before it (kernels enqueued ahead of
    it: a pre-pass, a kernel on another stream) to lead_ns before the next launch's mark."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end from kernels order by start").fetchall()
    rows = [r for r in rows if any(sub in r[0] for sub in subs)]
    marks = [r[1] for r in rows if subs[0] in r[0]]
    spans = []
    for i, m0 in enumerate(marks):
        m1 = marks[i + 1] - lead_ns if i + 1 < len(marks) else float("inf")
        sel = [r for r in rows if m0 - lead_ns <= r[1] < m1]
        spans.append(max(r[2] for r in sel) - min(r[1] for r in sel))
    return (sum(spans) / len(spans) if spans else 0.0), len(spans)


def counter_means(db: str, kernel_filter: str = ""):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, dispatch_id, counter_name, value, "
                       "start, end from counters_collection").fetchall()
    acc = defaultdict(list)
    for name, disp, cname, val, t0, t1 in rows:
        if kernel_filter in name:
            acc[(name, cname)].append(val)
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


PEAK_VALU_WAVE_INSTR_PER_S = 256 * 2.4e9      # one VALU wave-instruction per CU per clock

# trace_kernel<mode, counting, precision> names of the uncounted grid kernels
TRACE_KERNELS = {"fp64": "trace_kernel<0, false, 0>", "fp32_filter": "trace_kernel<0, false, 2>",
                 "fp32_trace": "trace_kernel<0, false, 1>"}


def traffic_summary(kt_db, fetch_db, write_db, cfetch_db, cwrite_db, sq_db, copy_mib: float,
                    paths: int, workload: str, kernels=None, tcc_db=None):
    copy_bytes = copy_mib * 2 ** 20

    def one(db, counter, sub):
        for (name, cname), (mean, _n) in counter_means(db, sub).items():
            if cname == counter:
                return name, mean
        raise KeyError("%s of %s not in %s" % (counter, sub, db))

    def per_launch(db, counter, subs):
        # the counter summed over every dispatch of the launch's kernels, per launch
        tot, nl = 0.0, 0
        for sub in subs:
            for (name, cname), (mean, n) in counter_means(db, sub).items():
                if cname == counter:
                    tot += mean * n
                    if sub == subs[0]:
                        nl = n
        if nl == 0:
            raise KeyError("%s of %s not in %s" % (counter, subs[0], db))
        return tot / nl

    _, cf = one(cfetch_db, "FETCH_SIZE", "copyBuffer")
    _, cw = one(cwrite_db, "WRITE_SIZE", "copyBuffer")
    fscale, wscale = copy_bytes / (cf * 1024.0), copy_bytes / (cw * 1024.0)
    kt = {k["name"]: k for k in kernel_stats(kt_db)}
    out = {"workload": workload, "fetch_scale": fscale, "write_scale": wscale,
           "peak_valu_wave_instr_per_s": PEAK_VALU_WAVE_INSTR_PER_S, "kernels": {}}
    for key, subs in (kernels or TRACE_KERNELS).items():
        sl = subs.split("+")
        names = [next(n for n in kt if sub in n) for sub in sl]
        hit = miss = 0.0
        if len(sl) == 1:
            avg_us = kt[names[0]]["average_ns"] / 1e3
            f = one(fetch_db, "FETCH_SIZE", sl[0])[1]
            w = one(write_db, "WRITE_SIZE", sl[0])[1]
            v = one(sq_db, "SQ_INSTS_VALU", sl[0])[1]
            if tcc_db:
                hit = one(tcc_db, "TCC_HIT_sum", sl[0])[1]
                miss = one(tcc_db, "TCC_MISS_sum", sl[0])[1]
        else:
            avg_us = launch_spans(kt_db, sl)[0] / 1e3
            f = per_launch(fetch_db, "FETCH_SIZE", sl)
            w = per_launch(write_db, "WRITE_SIZE", sl)
            v = per_launch(sq_db, "SQ_INSTS_VALU", sl)
            if tcc_db:
                hit = per_launch(tcc_db, "TCC_HIT_sum", sl)
                miss = per_launch(tcc_db, "TCC_MISS_sum", sl)
        fb, wb = f * 1024.0 * fscale, w * 1024.0 * wscale
        out["kernels"][key] = {
            "kernel": " + ".join(names), "average_us": avg_us,
            "hbm_fetch_bytes": fb, "hbm_write_bytes": wb, "hbm_bytes_per_launch": fb + wb,
            "valu_wave_instr_per_launch": v, "valu_wave_instr_per_path": v / paths,
            "valu_issue_fraction": v / (avg_us * 1e-6) / PEAK_VALU_WAVE_INSTR_PER_S}
        if tcc_db:
            out["kernels"][key]["l2_hit_rate"] = hit / max(1.0, hit + miss)
    return out


def main(argv):
    if len(argv) >= 11 and argv[1] == "traffic":
        import json
        kv = dict(a.split("=", 1) for a in argv[11:])
        tcc = kv.pop("tcc", None)
        print(json.dumps(traffic_summary(*argv[2:8], float(argv[8]), int(argv[9]), argv[10],
                                         kv or None, tcc), indent=1))
        return 0
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
