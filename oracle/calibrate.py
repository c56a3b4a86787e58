"""CPU-baseline calibration -- TEST INFRASTRUCTURE ONLY (development container).

The benchmark's CPU baseline on the GPU host is the oracle (the C restatement of
calcTransmit, liboracle.so): the reference itself may not leave this container.  This
script times, on identical inputs in this container, the UNMODIFIED reference
(src/RTMCphon.f90 with its own RANDOM_NUMBER and OpenMP loop, oracle/_ref/time_ref, built
by `make -C oracle ref timing`) and the oracle (one cell per task on host threads, as the
benchmark runs it) at 1, 2, 4 and 8 threads, and prints the paths/s of both and their
ratio: a baseline quoted in oracle paths/s converts to reference-equivalent paths/s by
that ratio.

    python oracle/calibrate.py [--out oracle/calibration.txt]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
import tempfile
import threading
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:] = [p for p in sys.path if os.path.abspath(p or ".") != HERE]   # the package, not oracle.py
for p in (os.path.join(ROOT, "p-trans_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import oracle as O  # noqa: E402
from ptrans import anphon, gen, geometry, mcp  # noqa: E402

TIME_REF = os.path.join(HERE, "_ref", "time_ref")

# name -> (model, sampling mean free path); bulk table 2 branches x 40 bins
CASES = {
    "C1 wire 1000x100x100 nm, p = 0 walls, ballistic":
        (lambda: gen.nanowire(p_wall=0.0, nphonon=400, nangle=11), 1.0e10),
    "C2 film 600x600x200 nm, 3 holes r = 50 nm, p = 0.3 (mirror y walls)":
        (lambda: gen.film_holes(seed=2, lx=600.0, ly=600.0, h=200.0, radius=50.0, nholes=3,
                                p_hole=0.3, p_film=0.0, p_side=1.0, nphonon=400, nangle=11), 1.0e10),
    "C3 cube 1000 nm, 8 spheres + 4 z-cylinders, Gamma law, Lambda = 200 nm":
        (lambda: gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.5, nphonon=150, nangle=11),
         200.0),
}


def reference_rate(tmp, threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([TIME_REF, "input.mcp", repr(tmp[1])], cwd=tmp[0], env=env,
                         capture_output=True, text=True, timeout=3600, check=True).stdout
    line = next(ln for ln in out.splitlines() if ln.startswith("TIME"))
    t = line.split()
    return int(t[3]) / float(t[1])


def oracle_rate(gdesc, mdesc, mfp, nangle, np_, threads):
    ncell = mdesc.nbranch * mdesc.nkpoint * (nangle - 1)
    lock = threading.Lock()
    state = {"next": 0, "paths": 0}

    def worker():
        while True:
            with lock:
                c = state["next"]
                state["next"] += 1
            if c >= ncell:
                return
            _, st = O.trace_grid(gdesc, mdesc, mfp, nangle, np_, 1, O.MATH_LIBM,
                                 cell_begin=c, cell_end=c + 1)
            with lock:
                state["paths"] += int(st[0])
    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker) for _ in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    return state["paths"] / (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--threads", default="1,2,4,8")
    args = ap.parse_args()
    if not os.path.exists(TIME_REF):
        sys.exit("oracle/_ref/time_ref is not built: make -C oracle ref timing")
    threads = [int(x) for x in args.threads.split(",")]
    table = anphon.synthetic_table(2, 40, 7)
    lines = ["# CPU calibration: unmodified reference (flang -O2 -fopenmp, own RANDOM_NUMBER) vs "
             "the oracle (gcc -O2, one cell per task), same inputs, this container (%d CPUs)"
             % os.cpu_count(),
             "# %-72s %7s %14s %14s %7s" % ("input", "threads", "reference p/s", "oracle p/s",
                                           "ratio")]
    for name, (factory, mfp) in CASES.items():
        model = factory()
        with tempfile.TemporaryDirectory() as d:
            os.makedirs(os.path.join(d, "Materials"))
            mcp.write_mcp(os.path.join(d, "input.mcp"), model)
            anphon.write_anphon(os.path.join(d, "Materials", "Si.anphon"), table)
            tab = anphon.read_anphon(os.path.join(d, "Materials", "Si.anphon"), "Si")
            geo = geometry.build_geometry(mcp.read_mcp(os.path.join(d, "input.mcp")))
            gdesc, _ga = geo.desc()
            mdesc, _ma = geometry.modes_desc(tab.omega, tab.vel, tab.mfp, tab.weight)
            for th in threads:
                r = reference_rate((d, mfp), th)
                o = oracle_rate(gdesc, mdesc, mfp, model.nangle, model.nphonon, th)
                lines.append("  %-72s %7d %14.4g %14.4g %7.3f" % (name, th, r, o, o / r))
                print(lines[-1], flush=True)
    text = "\n".join(lines) + "\n"
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
