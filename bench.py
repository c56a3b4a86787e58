"""Throughput of the phonon Monte Carlo trace on one MI355X.

Workload (the thin-film configuration): a 600 x 600 x 200 nm film, transport along x,
three z-cylindrical holes of radius 50 nm (seeded, non-overlapping; specularity 0.3,
vacuum t = 0), diffuse film faces (p = 0), periodic in-plane side walls (y; the
reference's walls always reflect, so periodic walls are this library's extension and
the oracle's restatement of them); synthetic bulk table 6 branches x 2000 frequency
bins; 11 polar angles (10 traced); Np particles per (k, b, angle) cell chosen so one
step traces ~1e8 phonon paths; ballistic (Matthiessen) trace in FP64 exactly as
calcTransmit (src/RTMCphon.f90:1608-2414) apart from the periodic wrap.  One step = one pt_trace_grid call, i.e.
TransmissionFunction (RT:1544-1604) for the whole table; inputs are resident in HBM
before timing starts.

Prints one JSON line (metric, value = paths/s, roofline of the trace kernel and the
CPU baseline: the oracle's C restatement on the host's cores, on a bounded sample).

The "adjoint" object measures the second half of the metric on the inclusions
configuration (1000 nm cube, 8 spheres + 4 z-cylinders, each inclusion with its own
specularity; same bulk table): the alias-mode forward trace and the likelihood-ratio
adjoint (12 specularity + 6 branch-scale derivatives of kappa) over 1e8 paths, and
kappa-gradient evaluations/s = adjoint paths/s / the paths that make the gradient
2-sigma significant (|stderr| = |grad| / 2).

The "grains" object measures the polycrystal configuration: a 1000 nm cube cut into 5
Voronoi grains whose shared boundaries transmit with the Lorentzian law
t = 1/(1 + (omega/omega0)^2), omega0 = 2 pi f0, f0 ~ U(3, 10) THz per boundary; alias
mode over 1e9 paths with per-grain heat-flux tallies (each grain's share of kappa), FP32
trace against FP64.

The "batch" object measures the screening configuration: 256 jittered simple-cubic grain
structures (4^3 grains, 2e3 - 1e5 triangles each, log-uniform) traced together in one
graph-replayed persistent launch with BVH scans, the leaf size and launch tile picked by
the project's autotuner, per-structure likelihood-ratio scores of the mean specularity,
p50/p99 per-structure completion times and the per-structure kappa error against FP64.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "p-trans_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

from ptrans import gen, geometry  # noqa: E402
from ptrans import post  # noqa: E402
from ptrans.anphon import Anphon, synthetic_table  # noqa: E402
from ptrans.constants import M_TWOPI, PREC_FP32_FILTER, PREC_FP32_RETEST, PREC_FP64  # noqa: E402

FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector, published
METRIC = "phonon paths/s and κ-gradient evals/s on 1 MI355X at fixed κ std-error"


def workload(np_per_cell: int, nangle: int = 11, nb: int = 6, nk: int = 2000):
    model = gen.film_holes(seed=2, lx=600.0, ly=600.0, h=200.0, radius=50.0, nholes=3,
                           p_hole=0.3, p_film=0.0, p_side=1.0, nphonon=np_per_cell, nangle=nangle)
    geo = geometry.set_periodic(geometry.build_geometry(model), axes=(1,))
    t = synthetic_table(nb, nk, seed=20)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12          # RT:727
    table = Anphon(material="synthetic", rtemp=t.rtemp, vol=t.vol, omega=omega, tau=t.tau_ps,
                   vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    return model, geo, omega, table


def kappa_of(counts, model, table, np_per_cell):
    """kappa, its binomial standard error and the paths needed for a 1e-3 relative error,
    through effectiveMFP + Matthiessen as the reference's fileoutput (RT:2418-2755)."""
    ang = post.vangle(model.nangle)
    semfp = post.effective_mfp(counts / float(np_per_cell), ang, model.box[0])[..., None]
    res = post.kappa(semfp, table, model.temperature, [1.0e10], True, counts=counts,
                     np_=np_per_cell, angles=ang, boxx=model.box[0])
    paths = counts.shape[0] * counts.shape[1] * (counts.shape[2] - 1) * np_per_cell
    need = paths * (res.kappa_stderr / (1e-3 * res.kappa)) ** 2 if res.kappa > 0 else None
    return res.kappa, res.kappa_stderr, need


def cpu_baseline(geo, omega, mfp, nangle, np_per_cell, seed, seconds, threads):
    """The oracle (C restatement, single-threaded per call) on `threads` host threads
    (ctypes releases the GIL), each tracing whole cells; runs cells until `seconds`."""
    from oracle import oracle as O
    gdesc, _ga = geo.desc()
    mdesc, _ma = geometry.modes_desc(omega, omega, omega, omega)
    ncell = omega.size * (nangle - 1)
    lock = threading.Lock()
    state = {"next": 0, "paths": 0}
    t_end = time.perf_counter() + seconds

    def worker():
        while time.perf_counter() < t_end:
            with lock:
                c = state["next"]
                state["next"] += 1
            if c >= ncell:
                return
            _, st = O.trace_grid(gdesc, mdesc, mfp, nangle, np_per_cell, seed, O.MATH_OWN,
                                 cell_begin=c, cell_end=c + 1)
            with lock:
                state["paths"] += int(st[0])
    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker) for _ in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    return state["paths"] / dt, state["paths"], min(state["next"], ncell), dt


def cpu_rate(work, units, seconds: float, threads: int):
    """The oracle on `threads` host threads (ctypes releases the GIL): each thread takes
    the next unit (a callable returning its path count) until `seconds` have passed.
    Returns (paths/s, paths, units done, seconds)."""
    lock = threading.Lock()
    state = {"next": 0, "paths": 0}
    t_end = time.perf_counter() + seconds

    def worker():
        while time.perf_counter() < t_end:
            with lock:
                i = state["next"]
                state["next"] += 1
            if i >= len(units):
                return
            n = work(units[i])
            with lock:
                state["paths"] += n
    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker) for _ in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    return state["paths"] / dt, state["paths"], min(state["next"], len(units)), dt


def alias_weights(omega, vel, mfp, weight, temperature=300.0):
    """C v Lambda w per mode (column-major), the alias weights of the library's sampler."""
    cap = np.vectorize(lambda o: post.heat_capacity(float(o), temperature))(omega)
    return (cap * vel * mfp * weight).reshape(-1, order="F")


def cpu_leg_baseline(work, units, seconds, threads, what):
    rate, npaths, nunits, dt = cpu_rate(work, units, seconds, threads)
    return {"value": rate, "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": "%d paths (%s) in %.1f s: the oracle's C restatement, %d host threads"
                      % (npaths, what, dt, threads)}


def adjoint_leg(ctx, npaths: int, table_seed: int = 20, cpu=None):
    """C3 inclusions: forward alias trace (FP32 trace and FP64) and the LR adjoint."""
    from ptrans.constants import PREC_FP32_RETEST as FAST
    model = gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.5, nphonon=1, nangle=3)
    geo = geometry.build_geometry(model)
    group, ngroup = geometry.inclusion_groups(geo)
    pvals = np.random.default_rng(3).uniform(0.05, 0.95, ngroup)
    p0 = geometry.group_specularity(geo.flatten()["surfaces"]["specularity"], group, pvals)
    t = synthetic_table(6, 2000, seed=table_seed)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    g = ctx.geometry(geo)
    g.set_specularity(p0)
    m = ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    length = geo.box[0]
    ctx.set_seed(777)
    ctx.trace_alias(g, m, 1_000_000, precision=FAST)                       # warm-up
    ctx.adjoint_alias(g, m, 1_000_000, surface_group=group, ngroup=ngroup, precision=FAST)
    t0 = time.perf_counter()
    mp, mt, fst = ctx.trace_alias(g, m, npaths, precision=FAST)
    t_fwd = time.perf_counter() - t0
    k32 = post.alias_kappa(mp, mt, t.mfp_nm, length)
    ctx.set_seed(778)
    _, _, pilot, _ = ctx.adjoint_alias(g, m, 1_000_000, nbatch=1, precision=FAST)
    base = pilot[0, 1] / pilot[0, 0]
    ctx.set_seed(777)              # the forward run's paths: both launches end with the same
    t0 = time.perf_counter()       # longest path, so the ratio is the adjoint's own cost
    _, _, sums, ast = ctx.adjoint_alias(g, m, npaths, surface_group=group, ngroup=ngroup,
                                        baseline=base, nbatch=64, precision=FAST)
    t_adj = time.perf_counter() - t0
    gr = post.alias_gradients(sums, length, ngroup)
    ctx.set_seed(777)
    mp64, mt64, st64 = ctx.trace_alias(g, m, npaths, precision=PREC_FP64)
    k64 = post.alias_kappa(mp64, mt64, t.mfp_nm, length)
    gnorm = float(np.linalg.norm(gr.grad))
    snorm = float(np.linalg.norm(gr.grad_stderr))
    n2s = npaths * (snorm / (0.5 * gnorm)) ** 2 if gnorm > 0 else None
    adj_rate = npaths / t_adj
    out = {
        "workload": "1000 nm cube, 8 spheres + 4 z-cylinders (Gamma law, per-inclusion "
                    "specularity U(0.05, 0.95)), 6x2000 modes, alias mode, FP32 trace",
        "paths": npaths,
        "forward_paths_per_s": npaths / t_fwd,
        "forward_kernel_ms": fst["kernel_ms"],
        "adjoint_paths_per_s": adj_rate,
        "adjoint_kernel_ms": ast["kernel_ms"],
        "adjoint_cost_vs_forward": t_adj / t_fwd,
        "components": ngroup + 6,
        "kappa_ratio": gr.ratio, "kappa_ratio_stderr": gr.stderr,
        "kappa_ratio_fp64": k64.ratio, "kappa_ratio_fp64_stderr": k64.stderr,
        "kappa_rel_err_vs_fp64": abs(k32.ratio - k64.ratio) / k64.ratio,
        "kappa_diff_in_sigma": abs(k32.ratio - k64.ratio) / float(np.hypot(k32.stderr, k64.stderr)),
        "fp64_forward_paths_per_s": npaths / max(1e-9, st64["kernel_ms"] * 1e-3),
        "grad_norm": gnorm, "grad_stderr_norm": snorm,
        "paths_for_2sigma_gradient": n2s,
        "kappa_gradient_evals_per_s": adj_rate / n2s if n2s else None,
        "paths_for_1e-3_rel_stderr": npaths * (k32.stderr / (1e-3 * k32.ratio)) ** 2,
        "seconds_for_1e-3_rel_stderr": npaths * (k32.stderr / (1e-3 * k32.ratio)) ** 2 * t_fwd / npaths,
        "events_per_path": ast["events"] / max(1, ast["paths"]),
        "cpu_baseline": None,
    }
    if cpu:
        from oracle import oracle as O
        gdesc, _ga = geo.desc()
        mdesc, _ma = geometry.modes_desc(omega, t.vel, t.mfp_nm, t.weight)
        prob, alias = O.build_alias(alias_weights(omega, t.vel, t.mfp_nm, t.weight))
        gsurf = np.ascontiguousarray(group, dtype=np.int32)

        def work(first):
            O.adjoint_alias(gdesc, mdesc, t.mfp_nm, prob, alias, 777, first, 2000, gsurf, ngroup,
                            base, nbatch=1)
            return 2000
        out["cpu_baseline"] = cpu_leg_baseline(work, list(range(0, npaths, 2000)), *cpu,
                                               "adjoint paths of the same workload, 2000 per task")
    return out


def grains_leg(ctx, npaths: int, table_seed: int = 20, cpu=None):
    """C4 polycrystal with Lorentzian boundaries and per-grain flux tallies (see the
    module docstring)."""
    model = gen.voronoi_cube(seed=4, L=1000.0, ngrain=5, kind="Gamma", value=2.0, p=0.2,
                             nphonon=1, nangle=3)
    geo = geometry.build_geometry(model)
    w0 = geometry.set_lorentz_boundaries(geo, np.random.default_rng(41), f0_thz=(3.0, 10.0))
    t = synthetic_table(6, 2000, seed=table_seed)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    g = ctx.geometry(geo)
    m = ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    out = {}
    for name, prec in (("fp32", PREC_FP32_RETEST), ("fp64", PREC_FP64)):
        ctx.set_seed(2024)
        mp, mt, gt, st = ctx.trace_alias_grains(g, m, npaths, precision=prec)
        k = post.alias_kappa(mp, mt, t.mfp_nm, geo.box[0])
        out[name] = dict(k=k, shares=post.grain_kappa_shares(gt, k.paths), st=st)
        print("grains: %s %.1f ms" % (name, st["kernel_ms"]), file=sys.stderr, flush=True)
    g.close()
    a, b = out["fp32"], out["fp64"]
    sig = float(np.hypot(a["k"].stderr, b["k"].stderr))
    cpu_b = None
    if cpu:
        from oracle import oracle as O
        gdesc, _ga = geo.desc()
        mdesc, _ma = geometry.modes_desc(omega, t.vel, t.mfp_nm, t.weight)
        prob, alias = O.build_alias(alias_weights(omega, t.vel, t.mfp_nm, t.weight))
        ngt = len(geo.grains)

        def work(first):
            O.trace_alias_grains(gdesc, mdesc, t.mfp_nm, prob, alias, 2024, first, 5000, ngt)
            return 5000
        cpu_b = cpu_leg_baseline(work, list(range(0, npaths, 5000)), *cpu,
                                 "alias paths with grain tallies, 5000 per task")
    return {
        "cpu_baseline": cpu_b,
        "workload": "1000 nm cube, 5 Voronoi grains, Lorentzian grain-boundary transmission "
                    "(f0 ~ U(3, 10) THz, %d boundaries), p = 0.2, 6x2000 modes, alias mode, "
                    "per-grain heat-flux tallies, FP32 trace" % len(w0),
        "paths": int(a["k"].paths),
        "paths_per_s": a["k"].paths / (a["st"]["kernel_ms"] * 1e-3),
        "kernel_ms": a["st"]["kernel_ms"],
        "events_per_path": a["st"]["events"] / max(1, a["st"]["paths"]),
        "fp64_paths_per_s": b["k"].paths / (b["st"]["kernel_ms"] * 1e-3),
        "speedup_vs_fp64": b["st"]["kernel_ms"] / a["st"]["kernel_ms"],
        "kappa_ratio": a["k"].ratio, "kappa_ratio_stderr": a["k"].stderr,
        "kappa_ratio_fp64": b["k"].ratio,
        "kappa_rel_err_vs_fp64": abs(a["k"].ratio - b["k"].ratio) / b["k"].ratio,
        "kappa_diff_in_sigma": abs(a["k"].ratio - b["k"].ratio) / sig,
        "paths_for_1e-3_rel_stderr": a["k"].paths * (a["k"].stderr / (1e-3 * a["k"].ratio)) ** 2,
        "seconds_for_1e-3_rel_stderr": (a["k"].paths * (a["k"].stderr / (1e-3 * a["k"].ratio)) ** 2
                                        * a["st"]["kernel_ms"] * 1e-3 / a["k"].paths),
        "grain_shares": [round(float(v), 6) for v in a["shares"][1:]],
        "grain_shares_sum_vs_kappa": float(a["shares"].sum() / a["k"].ratio),
    }


def batch_leg(ctx, nstruct: int, paths_per_struct: int, table_seed: int = 20, cpu=None):
    """C5 screening batch (see the module docstring)."""
    from ptrans import autotune
    from ptrans.trace import Batch
    t0 = time.perf_counter()
    geos, tris = gen.c5_structures(n=nstruct)
    t_gen = time.perf_counter() - t0
    say = lambda *a: print("batch:", *a, file=sys.stderr, flush=True)   # noqa: E731
    say("%d structures, %d triangles, generated in %.1f s" % (nstruct, sum(tris), t_gen))
    t = synthetic_table(3, 1000, seed=table_seed)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    m = ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    tune = autotune.tune_batch(ctx, geos, m, probe_paths=max(1000, paths_per_struct // 20),
                               leaf_sizes=(2, 4, 8), chunks=(256, 1024), refills=(16,),
                               repeats=1, precision=PREC_FP32_RETEST,
                               log=lambda r: say("autotune", json.dumps(r)))
    t0 = time.perf_counter()
    b = Batch(ctx, geos, paths_per_struct, leaf_size=tune["leaf_size"], grid=64)
    t_build = time.perf_counter() - t0
    info = b.info()
    say("built", json.dumps(info), "in %.1f s" % t_build)
    b.set_specularity(np.full(nstruct, 0.3))
    ctx.set_seed(4242)
    b.trace(m, precision=PREC_FP32_RETEST, graph=True)                     # graph build
    ctx.set_seed(4243)
    pilot, _, _ = b.trace(m, precision=PREC_FP32_RETEST, graph=True)
    base = pilot[:, 1] / np.maximum(pilot[:, 0], 1.0)
    ctx.set_seed(4244)
    t0 = time.perf_counter()
    sums, fin, st = b.trace(m, precision=PREC_FP32_RETEST, baseline=base, graph=True)
    wall = time.perf_counter() - t0
    say("fp32 trace %.1f ms" % st["kernel_ms"])
    ctx.set_seed(4244)
    s64, fin64, st64 = b.trace(m, precision=PREC_FP64, graph=True)
    b.close()
    c = 0.75 * geos[0].box[0]
    n = sums[:, 0]
    k32 = c * sums[:, 1] / n
    k64 = c * s64[:, 1] / s64[:, 0]
    var32 = np.maximum(sums[:, 2] / n - (sums[:, 1] / n) ** 2, 0.0)
    var64 = np.maximum(s64[:, 2] / s64[:, 0] - (s64[:, 1] / s64[:, 0]) ** 2, 0.0)
    sig = c * np.sqrt(var32 / n + var64 / s64[:, 0])
    rel = np.abs(k32 - k64) / k64
    # pooled over structures: the mean relative difference and its standard error
    d = (k32 - k64) / k64
    pooled, pooled_se = float(d.mean()), float(d.std(ddof=1) / np.sqrt(len(d)))
    grad = c * sums[:, 3] / n
    gvar = np.maximum(sums[:, 4] / n - (sums[:, 3] / n) ** 2, 0.0)
    gse = c * np.sqrt(gvar / n)
    gnorm, snorm = float(np.linalg.norm(grad)), float(np.linalg.norm(gse))
    n2s = float(n.sum()) * (snorm / (0.5 * gnorm)) ** 2 if gnorm > 0 else None
    rate = st["paths"] / (st["kernel_ms"] * 1e-3)
    ks = c * np.sqrt(var32 / n)
    need = n * (ks / (1e-3 * k32)) ** 2
    cpu_b = None
    if cpu:
        from oracle import oracle as O
        mdesc, _ma = geometry.modes_desc(omega, t.vel, t.mfp_nm, t.weight)
        prob, alias = O.build_alias(alias_weights(omega, t.vel, t.mfp_nm, t.weight))
        poff = np.concatenate([[0], np.cumsum(np.full(nstruct, paths_per_struct))])
        descs = {s: geos[s].desc() for s in range(0, nstruct, 16)}     # 16 structures
        units = [(s, int(poff[s]) + k) for k in range(0, min(10000, paths_per_struct), 500)
                 for s in descs]

        def work(u):
            s, first = u
            O.trace_alias(descs[s][0], mdesc, t.mfp_nm, prob, alias, 4244, first, 500)
            return 500
        cpu_b = cpu_leg_baseline(work, units, *cpu, "16 structures (every 16th) x up to 1e4 "
                                 "paths each, brute-force candidate scans, 500 per task")
    return {
        "cpu_baseline": cpu_b,
        "workload": "%d jittered 4x4x4 simple-cubic grain structures (1000 nm cube, +-5 nm "
                    "jitter, Gamma law, p = 0.3), %d - %d triangles each, 3x1000 modes, "
                    "%d paths per structure, one graph-replayed persistent launch, FP32 trace"
                    % (nstruct, min(tris), max(tris), paths_per_struct),
        "structures": nstruct, "triangles": int(sum(tris)), "surfaces": info["surfaces"],
        "bvh_nodes": info["nodes"], "bvh_depth": info["depth"],
        "paths": int(st["paths"]), "kernel_ms": st["kernel_ms"], "wall_s": wall,
        "paths_per_s": st["paths"] / (st["kernel_ms"] * 1e-3),
        "events_per_path": st["events"] / max(1, st["paths"]),
        "completion_ms_p50": float(np.percentile(fin, 50)),
        "completion_ms_p99": float(np.percentile(fin, 99)),
        "fp64_paths_per_s": st64["paths"] / (st64["kernel_ms"] * 1e-3),
        "kappa_rel_err_vs_fp64_median": float(np.median(rel)),
        "kappa_rel_err_vs_fp64_max": float(np.max(rel)),
        "kappa_diff_sigma_max": float(np.max(np.abs(k32 - k64) / sig)),
        "kappa_rel_diff_vs_fp64_pooled": pooled, "kappa_rel_diff_vs_fp64_pooled_stderr": pooled_se,
        "paths_for_1e-3_rel_stderr_per_structure_median": float(np.median(need)),
        "seconds_for_1e-3_rel_stderr_all_structures": float(need.sum() / rate),
        "mean_specularity_grad_median": float(np.median(grad)),
        "mean_specularity_grad_stderr_median": float(np.median(gse)),
        "gradient_components": nstruct,
        "paths_for_2sigma_gradient": n2s,
        "kappa_gradient_evals_per_s": rate / n2s if n2s else None,
        "autotune": {k: v for k, v in tune.items() if k != "table"},
        "autotune_configs": len(tune["table"]),
        "host_generate_s": t_gen, "host_build_s": t_build,
    }


def roofline(cstats, kernel_ms, kernel_key):
    """VALU roofline of the trace kernel: algorithmic FP64 and FP32 operations of one
    launch (instrumented twin kernel, same paths) over the timed kernel's average duration
    (HIP events on the library's stream), against MI355X's FP64 vector peak (non-packed
    FP32 issues at the same rate).  From the committed rocprofv3 summary of the same
    kernel: HBM traffic per launch (FETCH_SIZE/WRITE_SIZE scaled by a device copy) and
    the VALU issue rate, i.e. wave-instructions per second over 256 CUs x 2.4 GHz."""
    ops = cstats["fp64_ops"]
    achieved = ops / (kernel_ms * 1e-3) / 1e12
    prof = {}
    tf = os.path.join(ROOT, "profiles", "trace_kernel_film_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            prof = json.load(f).get("kernels", {}).get(kernel_key, {})
    return {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": prof.get("hbm_bytes_per_launch"),
            "fp64_ops_per_launch": ops, "fp64_ops_per_path": ops / max(1, cstats["paths"]),
            "philox_blocks_per_path": cstats["philox_blocks"] / max(1, cstats["paths"]),
            "valu_issue_fraction_profiled": prof.get("valu_issue_fraction"),
            "valu_wave_instr_per_path_profiled": prof.get("valu_wave_instr_per_path"),
            "profiled_kernel_us": prof.get("average_us")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--np", type=int, default=834, help="particles per cell (834 -> 1.0e8 paths)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-paths", type=float, default=1e8,
                    help="paths of the inclusions adjoint leg (0: skip it)")
    ap.add_argument("--grain-paths", type=float, default=1e9,
                    help="paths of the Voronoi/Lorentzian grains leg (0: skip it)")
    ap.add_argument("--batch-structures", type=int, default=256,
                    help="structures of the screening batch leg (0: skip it)")
    ap.add_argument("--batch-paths", type=float, default=1e5,
                    help="paths per structure of the screening batch leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (for HBM counter scaling)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.empty(args.calibrate_copy * 2 ** 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel()}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(paths would shard by id range, one range per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch
    from ptrans.trace import Context

    nangle, mfp, seed = 11, 1.0e10, 12345
    model, geo, omega, table = workload(args.np, nangle)
    nb, nk = omega.shape
    paths = nb * nk * (nangle - 1) * args.np
    res = {}
    with Context(seed=seed, device=0) as ctx:
        g = ctx.geometry(geo)
        m = ctx.modes(omega=omega)
        for prec in (PREC_FP64, PREC_FP32_FILTER, PREC_FP32_RETEST):
            for _ in range(args.warmup):
                ctx.trace_grid(g, m, mfp, nangle, args.np, precision=prec)
            torch.cuda.synchronize()
            kern_ms = []
            t0 = time.perf_counter()
            for _ in range(args.steps):
                counts, stats = ctx.trace_grid(g, m, mfp, nangle, args.np, precision=prec)
                kern_ms.append(stats["kernel_ms"])
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            # one more launch of the instrumented twin kernel (untimed) for the op count
            ctx.set_op_counting(True)
            _, cstats = ctx.trace_grid(g, m, mfp, nangle, args.np, precision=prec)
            ctx.set_op_counting(False)
            res[prec] = dict(dt=dt, kernel_ms=float(np.mean(kern_ms)), counts=counts, stats=stats,
                             cstats=cstats)
            print("film: precision %d, %.2f ms per step" % (prec, dt / args.steps * 1e3),
                  file=sys.stderr, flush=True)
        leg_cpu = None
        if not args.no_cpu:
            leg_cpu = (min(args.cpu_seconds, 10.0), args.cpu_threads or min(16, os.cpu_count() or 1))
        adj = (adjoint_leg(ctx, int(args.adjoint_paths), cpu=leg_cpu)
               if args.adjoint_paths > 0 else None)
        if adj:
            print("adjoint: %.1f ms" % adj["adjoint_kernel_ms"], file=sys.stderr, flush=True)
        grn = grains_leg(ctx, int(args.grain_paths), cpu=leg_cpu) if args.grain_paths > 0 else None
        bat = (batch_leg(ctx, args.batch_structures, int(args.batch_paths), cpu=leg_cpu)
               if args.batch_structures > 0 else None)
    fast, plain, filt = res[PREC_FP32_RETEST], res[PREC_FP64], res[PREC_FP32_FILTER]
    dt = fast["dt"]
    value = paths * args.steps / dt
    k_ms = fast["kernel_ms"]
    cstats = fast["cstats"]
    stats = fast["stats"]
    counts = fast["counts"]
    events_per_path = stats["events"] / stats["paths"]
    transmitted = counts.sum() / paths
    kap, kap_err, need = kappa_of(counts, model, table, args.np)
    kap64, _, _ = kappa_of(plain["counts"], model, table, args.np)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        rate, npaths, ncells, cdt = cpu_baseline(geo, omega, mfp, nangle, args.np, seed,
                                                 args.cpu_seconds, threads)
        cpu = {"value": rate, "unit": "paths/s", "cores": threads, "kind": "port",
               "sample": "%d whole cells (%d paths) of the same workload in %.1f s: the oracle's "
                         "C restatement, %d host threads, one cell per call"
                         % (ncells, npaths, cdt, threads)}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded film geometry and bulk table)",
        "config": {"workload": "thin film 600x600x200 nm, 3 z-cylinder holes r=50 nm p=0.3, "
                               "periodic in-plane (y) side walls, "
                               "6x2000 modes x 10 angles x %d particles, ballistic grid "
                               "trace (TransmissionFunction)" % args.np,
                   "paths_per_step": paths, "events_per_path": events_per_path,
                   "transmission": transmitted},
        "kernel_ms": k_ms,
        "precision": "FP32 intersection filter and FP32 direction sampling/reflection; FP64 "
                     "retests, hit distance, positions, free paths and tallies",
        "fp32_filter_bitexact": {"value": paths * args.steps / filt["dt"],
                                 "kernel_ms": filt["kernel_ms"],
                                 "counts_identical_to_fp64": bool(np.array_equal(filt["counts"], plain["counts"]))},
        "fp64_plain": {"value": paths * args.steps / plain["dt"],
                       "ms_per_step": plain["dt"] / args.steps * 1e3,
                       "kernel_ms": plain["kernel_ms"],
                       "roofline": roofline(plain["cstats"], plain["kernel_ms"], "fp64")},
        "speedup_vs_fp64": plain["dt"] / dt,
        "kappa": {"value_W_mK": kap, "stderr": kap_err, "rel_err_vs_fp64": abs(kap - kap64) / kap64,
                  "counts_identical_to_fp64": bool(np.array_equal(counts, plain["counts"])),
                  "paths_for_1e-3_rel_stderr": need,
                  "seconds_for_1e-3_rel_stderr": need / value if need else None},
        "retests_per_event": stats["retests"] / max(1, stats["events"]),
        "fallbacks_per_event": stats["fallbacks"] / max(1, stats["events"]),
        "roofline": roofline(cstats, k_ms, "fp32_trace"),
        "cpu_baseline": cpu,
        "adjoint": adj,
        "grains": grn,
        "batch": bat,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
