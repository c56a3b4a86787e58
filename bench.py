"""Throughput of the phonon Monte Carlo trace on one MI355X.

Headline workload (the polycrystal configuration, the largest one that runs at its stated
size on one GPU within a bench step): a 1000 nm cube cut into 5 Voronoi grains (seeded
nucleation points, planar facets by half-space clipping), every grain boundary transmitting
with the Lorentzian law t = 1/(1 + (omega/omega0)^2), omega0 = 2 pi f0 with f0 ~ U(3, 10) THz
per boundary, specularity 0.2, diffuse box walls; synthetic bulk table of 6 branches x 2000
bins (seeded).  A step is one alias-mode trace of 1e9 phonon paths: mode sampling from the
device alias table (weight C v Lambda w), Lambertian incidence on x = 0, per-mode exponential
free flights, the reference's event loop (calcTransmit, src/RTMCphon.f90:1693-1824) with
specular/diffuse scattering and grain-boundary transmission, per-mode integer tallies and
per-grain FP64 heat-flux tallies.  Precision PT_PREC_FP32_FILTER: FP32 intersection bounds
with FP64 certification, every result computed in FP64 -- bit-identical to the FP64 trace.
Inputs are resident in HBM before timing starts.

Prints one JSON line: metric, value = paths/s over the timed steps, the roofline of the
trace kernel (algorithmic FP64 operations per path from the op-counting kernel, over the
kernel's average duration from HIP events on the library's stream, against MI355X's FP64
peak; HBM traffic per launch from the committed rocprofv3 counter summary), and the CPU
baseline (the oracle's C restatement on all of the host's cores, bounded sample).

Other legs of the same line (reported, not `value`):
  film     the thin-film configuration: 600 x 600 x 200 nm, 3 z-cylinder holes r = 50 nm,
           p = 0.3, periodic y walls; 6 x 2000 modes x 10 angles x 834 particles (1e8 paths)
           through pt_trace_grid (TransmissionFunction) in all three precisions;
  adjoint  the inclusions configuration: 1000 nm cube, 8 spheres + 4 z-cylinders, 1e8 paths
           forward and likelihood-ratio adjoint (12 specularity + 6 branch-scale derivatives
           of kappa), kappa-gradient evaluations/s;
  batch    the screening configuration: 256 jittered 4 x 4 x 4 simple-cubic grain structures
           of 2e3 - 1e5 triangles, one persistent launch per replay (fresh keys) with BVH
           scans and 32 live tail workgroups beside it, long walks handed over after 200
           events (--batch-graph: HIP-graph replays, whose tail kernel runs after the
           persistent one), tile and leaf size from the autotuner, p50/p99 per-structure
           completion, per-structure kappa against the bit-identical precision.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
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

FP64_PEAK_TFLOPS = 78.6     # MI355X, published (vector and matrix FP64)
FP32_PEAK_TFLOPS = 157.3    # MI355X, published FP32 (the 2x rate needs packed FP32)
HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, published
METRIC = "phonon paths/s and κ-gradient evals/s on 1 MI355X at fixed κ std-error"
PROFILES = os.path.join(ROOT, "profiles")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------------------ host CPU

def host_cpu():
    """Cores the baseline uses (every CPU this process may run on) and what lscpu says."""
    info = {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU family": "cpu_family",
                "Model": "cpu_model_number", "Hypervisor vendor": "hypervisor"}
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


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


def cpu_leg_baseline(work, units, cpu, what):
    seconds, threads = cpu
    rate, npaths, nunits, dt = cpu_rate(work, units, seconds, threads)
    return {"value": rate, "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": "%d paths (%s) in %.1f s: the oracle's C restatement (single-threaded per "
                      "task, as the reference's OpenMP loop runs one cell per thread) on %d host "
                      "threads = every CPU of the host" % (npaths, what, dt, threads)}


def alias_weights(omega, vel, mfp, weight, temperature=300.0):
    """C v Lambda w per mode (column-major), the alias weights of the library's sampler."""
    cap = np.vectorize(lambda o: post.heat_capacity(float(o), temperature))(omega)
    return (cap * vel * mfp * weight).reshape(-1, order="F")


# ------------------------------------------------------------------------------ roofline

def load_profile(name):
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def roofline(ops_per_path, paths, kernel_ms, peak, prof=None, extra=None):
    """VALU roofline (the trace is neither a contraction nor bandwidth-bound): algorithmic
    operations per path -- the FP64 operations of the reference's arithmetic, counted by
    the instrumented kernel on the FP64 trace of the same workload -- times the paths of a
    launch, over the launch's average kernel time (HIP events on the library's stream);
    HBM traffic per launch and VALU issue from the committed rocprofv3 counter summary."""
    prof = prof or {}
    achieved = ops_per_path * paths / (kernel_ms * 1e-3) / 1e12
    out = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
           "frac": achieved / peak, "traffic": prof.get("hbm_bytes_per_launch"),
           "ops_per_path": ops_per_path, "paths_per_launch": paths, "kernel_ms": kernel_ms,
           "valu_issue_fraction_profiled": prof.get("valu_issue_fraction"),
           "valu_wave_instr_per_path_profiled": prof.get("valu_wave_instr_per_path"),
           "profiled_kernel_us": prof.get("average_us")}
    if extra:
        out.update(extra)
    return out


def counted_ops(st):
    return st["fp64_ops"] / max(1, st["paths"]), st["philox_blocks"] / max(1, st["paths"])


# ------------------------------------------------------------------------------ headline

def polycrystal(ctx, table_seed=20):
    model = gen.voronoi_cube(seed=4, L=1000.0, ngrain=5, kind="Gamma", value=2.0, p=0.2,
                             nphonon=1, nangle=3)
    geo = geometry.build_geometry(model)
    w0 = geometry.set_lorentz_boundaries(geo, np.random.default_rng(41), f0_thz=(3.0, 10.0))
    t = synthetic_table(6, 2000, seed=table_seed)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    g = ctx.geometry(geo)
    m = ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    return geo, w0, t, omega, g, m


def headline(ctx, npaths: int, steps: int, warmup: int, count_paths: int, cmp_paths: int, cpu):
    geo, w0, t, omega, g, m = polycrystal(ctx)
    seed0 = 2024
    for k in range(warmup):
        ctx.set_seed(seed0 - 1 - k)
        ctx.trace_alias_grains(g, m, npaths, precision=PREC_FP32_FILTER)
    kern, res = [], None
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.set_seed(seed0 + k)
        mp, mt, gt, st = ctx.trace_alias_grains(g, m, npaths, precision=PREC_FP32_FILTER)
        kern.append(st["kernel_ms"])
        if k == 0:
            res = (mp, mt, gt, st)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    say("headline: %d steps, %.1f ms per step" % (steps, dt / steps * 1e3))
    mp, mt, gt, st = res
    k_f = post.alias_kappa(mp, mt, t.mfp_nm, geo.box[0])
    shares = post.grain_kappa_shares(gt, k_f.paths)
    # one set of paths (cmp_paths, a fresh key) in the three precisions: the filter's
    # tallies equal FP64's, the FP32 trace's kappa is compared with FP64's
    cseed = seed0 + 10000
    cmp = {}
    for prec in ((PREC_FP32_FILTER, PREC_FP64, PREC_FP32_RETEST) if cmp_paths > 0 else ()):
        ctx.set_seed(cseed)
        cmp[prec] = ctx.trace_alias_grains(g, m, cmp_paths, precision=prec)
        say("headline: %.3g paths in precision %d, %.1f ms" % (cmp_paths, prec, cmp[prec][3]["kernel_ms"]))
    if cmp_paths <= 0:          # profiling runs: the headline launches only
        cmp = {p: (mp, mt, gt, st) for p in (PREC_FP32_FILTER, PREC_FP64, PREC_FP32_RETEST)}
        cmp_paths = npaths
    mpf, mtf, _, stf = cmp[PREC_FP32_FILTER]
    mp64, mt64, _, st64 = cmp[PREC_FP64]
    mp32, mt32, _, st32 = cmp[PREC_FP32_RETEST]
    k_64 = post.alias_kappa(mp64, mt64, t.mfp_nm, geo.box[0])
    k_32 = post.alias_kappa(mp32, mt32, t.mfp_nm, geo.box[0])
    # algorithmic operations per path: the instrumented FP64 kernel on a sample of the paths
    ctx.set_seed(seed0)
    ctx.set_op_counting(True)
    _, _, _, cst = ctx.trace_alias_grains(g, m, count_paths, precision=PREC_FP64)
    ctx.set_op_counting(False)
    ops, blocks = counted_ops(cst)
    kms = float(np.mean(kern))
    prof = load_profile("trace_kernel_polycrystal_traffic.json").get("kernels", {})
    sig = float(np.hypot(k_32.stderr, k_64.stderr))
    out = {
        "workload": "polycrystal 1000 nm cube, 5 Voronoi grains, Lorentzian grain boundaries "
                    "t = 1/(1 + (w/w0)^2), f0 ~ U(3, 10) THz (%d boundaries), p = 0.2, "
                    "synthetic 6x2000 bulk table, alias-mode sampling, Lambertian incidence, "
                    "per-mode free flights, per-grain heat-flux tallies, %.3g paths per step"
                    % (len(w0), npaths),
        "dt": dt, "kernel_ms": kms, "kernel_ms_per_step": kern,
        "events_per_path": st["events"] / max(1, st["paths"]),
        "kappa": {"ratio_to_bulk": k_f.ratio, "stderr": k_f.stderr,
                  "tallies_identical_to_fp64": bool(np.array_equal(mpf, mp64) and np.array_equal(mtf, mt64)),
                  "compared_paths": cmp_paths,
                  "paths_for_1e-3_rel_stderr": k_f.paths_for_rel_error(1e-3),
                  "seconds_for_1e-3_rel_stderr": k_f.paths_for_rel_error(1e-3) / (npaths / (kms * 1e-3)),
                  "grain_shares": [round(float(v), 6) for v in shares[1:]],
                  "grain_shares_sum_vs_kappa": float(shares.sum() / k_f.ratio)},
        "same_paths_in_each_precision": {"paths": cmp_paths, "fp32_filter_kernel_ms": stf["kernel_ms"]},
        "fp64_plain": {"paths_per_s": cmp_paths / (st64["kernel_ms"] * 1e-3), "kernel_ms": st64["kernel_ms"],
                       "roofline": roofline(ops, cmp_paths, st64["kernel_ms"], FP64_PEAK_TFLOPS,
                                            prof.get("fp64"))},
        "fp32_trace": {"paths_per_s": cmp_paths / (st32["kernel_ms"] * 1e-3), "kernel_ms": st32["kernel_ms"],
                       "kappa_ratio": k_32.ratio, "kappa_rel_err_vs_fp64": abs(k_32.ratio - k_64.ratio) / k_64.ratio,
                       "kappa_diff_in_sigma": abs(k_32.ratio - k_64.ratio) / sig,
                       "roofline": roofline(ops, cmp_paths, st32["kernel_ms"], FP32_PEAK_TFLOPS,
                                            prof.get("fp32_trace"),
                                            {"frac_vs_fp64_peak_78_6": ops * cmp_paths / (st32["kernel_ms"] * 1e-3)
                                             / 1e12 / FP64_PEAK_TFLOPS})},
        "speedup_vs_fp64_same_paths": st64["kernel_ms"] / stf["kernel_ms"],
        "fp32_trace_speedup_vs_fp64_same_paths": st64["kernel_ms"] / st32["kernel_ms"],
        "roofline": roofline(ops, npaths, kms, FP64_PEAK_TFLOPS, prof.get("fp32_filter"),
                             {"philox_blocks_per_path": blocks, "op_count_sample_paths": count_paths}),
        "cpu_baseline": None,
    }
    if cpu:
        from oracle import oracle as O
        gdesc, _ga = geo.desc()
        mdesc, _ma = geometry.modes_desc(omega, t.vel, t.mfp_nm, t.weight)
        prob, alias = O.build_alias(alias_weights(omega, t.vel, t.mfp_nm, t.weight))
        ngt = len(geo.grains)

        def work(first):
            O.trace_alias_grains(gdesc, mdesc, t.mfp_nm, prob, alias, seed0, first, 5000, ngt)
            return 5000
        out["cpu_baseline"] = cpu_leg_baseline(work, list(range(0, npaths, 5000)), cpu,
                                               "alias paths of the same workload with grain "
                                               "tallies, FP64, 5000 per task")
    g.close()
    m.close()
    return out


# ------------------------------------------------------------------------------ film leg

def film_workload(np_per_cell: int, nangle: int = 11, nb: int = 6, nk: int = 2000):
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


def film_leg(ctx, np_per_cell: int, steps: int, warmup: int, cpu):
    import torch
    nangle, mfp, seed = 11, 1.0e10, 12345
    model, geo, omega, table = film_workload(np_per_cell, nangle)
    nb, nk = omega.shape
    paths = nb * nk * (nangle - 1) * np_per_cell
    g = ctx.geometry(geo)
    m = ctx.modes(omega=omega)
    ctx.set_seed(seed)
    res = {}
    for prec in (PREC_FP64, PREC_FP32_FILTER, PREC_FP32_RETEST):
        for _ in range(warmup):
            ctx.trace_grid(g, m, mfp, nangle, np_per_cell, precision=prec)
        torch.cuda.synchronize()
        kern = []
        t0 = time.perf_counter()
        for _ in range(steps):
            counts, stats = ctx.trace_grid(g, m, mfp, nangle, np_per_cell, precision=prec)
            kern.append(stats["kernel_ms"])
        torch.cuda.synchronize()
        res[prec] = dict(dt=time.perf_counter() - t0, kernel_ms=float(np.mean(kern)),
                         counts=counts, stats=stats)
        say("film: precision %d, %.2f ms per step" % (prec, res[prec]["dt"] / steps * 1e3))
    ctx.set_op_counting(True)
    _, cst = ctx.trace_grid(g, m, mfp, nangle, np_per_cell, precision=PREC_FP64)
    ctx.set_op_counting(False)
    ops, blocks = counted_ops(cst)
    prof = load_profile("trace_kernel_film_traffic.json").get("kernels", {})
    f64, flt, f32 = res[PREC_FP64], res[PREC_FP32_FILTER], res[PREC_FP32_RETEST]
    kap, kap_err, need = kappa_of(f32["counts"], model, table, np_per_cell)
    kap64, _, _ = kappa_of(f64["counts"], model, table, np_per_cell)
    rate = lambda r: paths * steps / r["dt"]      # noqa: E731
    out = {
        "workload": "thin film 600x600x200 nm, 3 z-cylinder holes r=50 nm p=0.3, periodic "
                    "in-plane (y) side walls, 6x2000 modes x 10 angles x %d particles, "
                    "ballistic grid trace (TransmissionFunction)" % np_per_cell,
        "paths_per_step": paths,
        "events_per_path": f32["stats"]["events"] / f32["stats"]["paths"],
        "transmission": float(f32["counts"].sum() / paths),
        "fp64_plain": {"paths_per_s": rate(f64), "kernel_ms": f64["kernel_ms"],
                       "roofline": roofline(ops, paths, f64["kernel_ms"], FP64_PEAK_TFLOPS,
                                            prof.get("fp64"), {"philox_blocks_per_path": blocks})},
        "fp32_filter_bitexact": {"paths_per_s": rate(flt), "kernel_ms": flt["kernel_ms"],
                                 "counts_identical_to_fp64": bool(np.array_equal(flt["counts"], f64["counts"])),
                                 "roofline": roofline(ops, paths, flt["kernel_ms"], FP64_PEAK_TFLOPS,
                                                      prof.get("fp32_filter"))},
        "fp32_trace": {"paths_per_s": rate(f32), "kernel_ms": f32["kernel_ms"],
                       "roofline": roofline(ops, paths, f32["kernel_ms"], FP32_PEAK_TFLOPS,
                                            prof.get("fp32_trace"),
                                            {"frac_vs_fp64_peak_78_6": ops * paths / (f32["kernel_ms"] * 1e-3)
                                             / 1e12 / FP64_PEAK_TFLOPS})},
        "speedup_fp32_trace_vs_fp64": f64["dt"] / f32["dt"],
        "kappa": {"value_W_mK": kap, "stderr": kap_err, "rel_err_vs_fp64": abs(kap - kap64) / kap64,
                  "paths_for_1e-3_rel_stderr": need,
                  "seconds_for_1e-3_rel_stderr": need / rate(f32) if need else None},
        "retests_per_event": f32["stats"]["retests"] / max(1, f32["stats"]["events"]),
        "fallbacks_per_event": f32["stats"]["fallbacks"] / max(1, f32["stats"]["events"]),
        "cpu_baseline": None,
    }
    if cpu:
        from oracle import oracle as O
        gdesc, _ga = geo.desc()
        mdesc, _ma = geometry.modes_desc(omega, omega, omega, omega)

        def work(c):
            _, st = O.trace_grid(gdesc, mdesc, mfp, nangle, np_per_cell, seed, O.MATH_OWN,
                                 cell_begin=c, cell_end=c + 1)
            return int(st[0])
        ncell = nb * nk * (nangle - 1)
        out["cpu_baseline"] = cpu_leg_baseline(work, list(range(ncell)), cpu,
                                               "whole (k, b, angle) cells of the same workload, FP64")
        out["fp32_trace"]["vs_cpu"] = out["fp32_trace"]["paths_per_s"] / out["cpu_baseline"]["value"]
        out["fp32_filter_bitexact"]["vs_cpu"] = (out["fp32_filter_bitexact"]["paths_per_s"]
                                                 / out["cpu_baseline"]["value"])
    g.close()
    m.close()
    return out


# ------------------------------------------------------------------------------ adjoint leg

def adjoint_leg(ctx, npaths: int, cpu, table_seed: int = 20):
    """C3 inclusions: forward alias trace (FP32 trace and FP64) and the LR adjoint."""
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
    FAST = PREC_FP32_RETEST
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
    # the same adjoint at ten times the paths per launch: the tail of the few longest random
    # walks (which sets a 1e8-path launch) amortised over more paths
    big = 10 * npaths
    ctx.set_seed(779)
    _, _, _, bst = ctx.adjoint_alias(g, m, big, surface_group=group, ngroup=ngroup,
                                     baseline=base, nbatch=64, precision=FAST)
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
        "adjoint_cost_vs_forward": ast["kernel_ms"] / fst["kernel_ms"],
        "adjoint_paths_per_s_at_10x_paths": big / (bst["kernel_ms"] * 1e-3),
        "adjoint_kernel_ms_at_10x_paths": bst["kernel_ms"],
        "adjoint_wall_vs_forward_wall": t_adj / t_fwd,
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
        out["cpu_baseline"] = cpu_leg_baseline(work, list(range(0, npaths, 2000)), cpu,
                                               "adjoint paths of the same workload, FP64, 2000 per task")
        out["adjoint_vs_cpu"] = adj_rate / out["cpu_baseline"]["value"]
        out["adjoint_vs_cpu_at_10x_paths"] = out["adjoint_paths_per_s_at_10x_paths"] / out["cpu_baseline"]["value"]
    g.close()
    m.close()
    return out


# ------------------------------------------------------------------------------ batch leg

def batch_leg(ctx, nstruct: int, paths_per_struct: int, replays: int, cpu, table_seed: int = 20,
              leaf: int = 0, graph: bool = False):
    """C5 screening batch: `replays` launches of paths_per_struct paths per structure (each
    replay a fresh Philox key; graph: HIP-graph replays), FP32 trace, and the same replays
    in the bit-identical precision (the FP64 results) for the per-structure kappa error.
    The tail kernel runs 32 live workgroups beside the persistent one (plain launches)."""
    from ptrans import autotune
    from ptrans.trace import Batch
    t0 = time.perf_counter()
    geos, tris = gen.c5_structures(n=nstruct)
    t_gen = time.perf_counter() - t0
    say("batch: %d structures, %d triangles, generated in %.1f s" % (nstruct, sum(tris), t_gen))
    t = synthetic_table(3, 1000, seed=table_seed)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    m = ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    if leaf > 0:            # fixed setting (profiling runs): no tuning launches
        ctx.set_tuning(256, 16, 0)
        tune = {"leaf_size": leaf, "top_levels": 1, "chunk": 256, "refill_min": 16,
                "blocks_per_cu": 0, "table": []}
    else:
        tune = autotune.tune_batch(ctx, geos, m, probe_paths=max(1000, paths_per_struct // 20),
                                   leaf_sizes=(2, 4, 8), chunks=(256, 1024), refills=(16,),
                                   repeats=1, precision=PREC_FP32_RETEST,
                                   log=lambda r: say("autotune", json.dumps(r)))
    t0 = time.perf_counter()
    b = Batch(ctx, geos, paths_per_struct, leaf_size=tune["leaf_size"], grid=64)
    b.set_top_levels(tune.get("top_levels", 0))
    t_build = time.perf_counter() - t0
    info = b.info()
    say("batch: built", json.dumps(info), "in %.1f s" % t_build)
    b.set_specularity(np.full(nstruct, 0.3))
    # mesh walks cost more per event in the persistent kernel (BVH scans, no safety balls)
    # than in the tail kernel: hand them over after 200 events, not 500 (3% faster launches
    # over five seeds; 32 live workgroups beat 16 and 48)
    ctx.set_tail(200, 1 << 20, 32)
    ctx.set_seed(4242)
    b.trace(m, precision=PREC_FP32_RETEST, graph=graph)                    # warm-up / graph build
    ctx.set_seed(4243)
    pilot, _, _ = b.trace(m, precision=PREC_FP32_RETEST, graph=graph)
    base = pilot[:, 1] / np.maximum(pilot[:, 0], 1.0)
    acc = {PREC_FP32_RETEST: np.zeros((nstruct, 5)), PREC_FP32_FILTER: np.zeros((nstruct, 5))}
    kern = {PREC_FP32_RETEST: [], PREC_FP32_FILTER: []}
    fins, walls = [], []
    for prec in (PREC_FP32_RETEST, PREC_FP32_FILTER):
        for r in range(replays):
            ctx.set_seed(5000 + r)
            t0 = time.perf_counter()
            sums, fin, st = b.trace(m, precision=prec, baseline=base, graph=graph)
            if prec == PREC_FP32_RETEST:
                walls.append(time.perf_counter() - t0)
                fins.append(fin)
            acc[prec] += sums
            kern[prec].append(st["kernel_ms"])
        say("batch: precision %d, %d replays, %.1f ms per replay" % (prec, replays, np.mean(kern[prec])))
    ctx.set_tail()
    # the FP64 kernel itself on every 16th structure (1e4 paths each): its sums equal the
    # filter's on the same paths, and its rate
    sub = geos[::16]
    bs = Batch(ctx, sub, 10000, leaf_size=tune["leaf_size"], grid=64)
    bs.set_top_levels(tune.get("top_levels", 0))
    bs.set_specularity(np.full(len(sub), 0.3))
    ctx.set_seed(5000)
    s64, _, st64 = bs.trace(m, precision=PREC_FP64)
    ctx.set_seed(5000)
    s_f1, _, _ = bs.trace(m, precision=PREC_FP32_FILTER)
    bs.close()
    ctx.set_op_counting(True)
    ctx.set_seed(5000)
    _, _, cst = b.trace(m, precision=PREC_FP32_RETEST)
    ctx.set_op_counting(False)
    b.close()
    sums, ref = acc[PREC_FP32_RETEST], acc[PREC_FP32_FILTER]
    c = 0.75 * geos[0].box[0]
    n, nr = sums[:, 0], ref[:, 0]
    k32, k64 = c * sums[:, 1] / n, c * ref[:, 1] / nr
    var32 = np.maximum(sums[:, 2] / n - (sums[:, 1] / n) ** 2, 0.0)
    var64 = np.maximum(ref[:, 2] / nr - (ref[:, 1] / nr) ** 2, 0.0)
    sig = c * np.sqrt(var32 / n + var64 / nr)
    rel = np.abs(k32 - k64) / k64
    d = (k32 - k64) / k64
    grad = c * sums[:, 3] / n
    gvar = np.maximum(sums[:, 4] / n - (sums[:, 3] / n) ** 2, 0.0)
    gse = c * np.sqrt(gvar / n)
    gnorm, snorm = float(np.linalg.norm(grad)), float(np.linalg.norm(gse))
    n2s = float(n.sum()) * (snorm / (0.5 * gnorm)) ** 2 if gnorm > 0 else None
    kms = float(np.mean(kern[PREC_FP32_RETEST]))
    paths_launch = nstruct * paths_per_struct
    rate = paths_launch / (kms * 1e-3)
    ks = c * np.sqrt(var32 / n)
    need = n * (ks / (1e-3 * k32)) ** 2
    fin = np.concatenate(fins)
    nodes, tests = cst["bvh_nodes"] / cst["paths"], cst["bvh_tests"] / cst["paths"]
    # algorithmic bytes per path: 32 B alias + mode record, 32 B per node fetch, 40 B per
    # triangle test
    alg_bytes = 32.0 + 32.0 * nodes + 40.0 * tests
    achieved_gbs = alg_bytes * paths_launch / (kms * 1e-3) / 1e9
    prof = load_profile("trace_kernel_batch_traffic.json").get("kernels", {}).get("fp32_trace", {})
    traffic = None
    if prof.get("hbm_bytes_per_launch") and prof.get("valu_wave_instr_per_path"):
        # the profile's launch may hold another path count: scale its bytes per path
        prof_paths = prof["valu_wave_instr_per_launch"] / prof["valu_wave_instr_per_path"]
        traffic = prof["hbm_bytes_per_launch"] / prof_paths * paths_launch
    out = {
        "workload": "%d jittered 4x4x4 simple-cubic grain structures (1000 nm cube, +-5 nm "
                    "jitter, Gamma law, p = 0.3), %d - %d triangles each, 3x1000 modes, "
                    "%d paths per structure per launch x %d %s, FP32 trace"
                    % (nstruct, min(tris), max(tris), paths_per_struct, replays,
                       "graph replays" if graph
                       else "launches (32 live tail workgroups, hand-over after 200 events)"),
        "structures": nstruct, "triangles": int(sum(tris)), "surfaces": info["surfaces"],
        "bvh_nodes": info["nodes"], "bvh_depth": info["depth"],
        "paths_per_structure": int(paths_per_struct * replays),
        "paths_per_launch": paths_launch, "kernel_ms": kms,
        "wall_s_per_launch": float(np.mean(walls)),
        "paths_per_s": rate,
        "events_per_path": cst["events"] / max(1, cst["paths"]),
        "completion_ms_p50": float(np.percentile(fin, 50)),
        "completion_ms_p99": float(np.percentile(fin, 99)),
        "fp32_filter_paths_per_s": paths_launch / (np.mean(kern[PREC_FP32_FILTER]) * 1e-3),
        "fp64_paths_per_s": st64["paths"] / (st64["kernel_ms"] * 1e-3),
        "fp64_sample": "%d structures (every 16th) x 1e4 paths" % len(sub),
        "fp64_sums_equal_filter_sums": bool(np.array_equal(s64[:, 0], s_f1[:, 0])
                                            and np.allclose(s64[:, 1], s_f1[:, 1], rtol=1e-12, atol=0)),
        "kappa_rel_err_vs_fp64_median": float(np.median(rel)),
        "kappa_rel_err_vs_fp64_p90": float(np.percentile(rel, 90)),
        "kappa_rel_err_vs_fp64_max": float(np.max(rel)),
        "structures_under_0.5pct": int(np.sum(rel < 5e-3)),
        "kappa_diff_sigma_max": float(np.max(np.abs(k32 - k64) / sig)),
        "kappa_rel_diff_vs_fp64_pooled": float(d.mean()),
        "kappa_rel_diff_vs_fp64_pooled_stderr": float(d.std(ddof=1) / np.sqrt(len(d))),
        "paths_for_1e-3_rel_stderr_per_structure_median": float(np.median(need)),
        "seconds_for_1e-3_rel_stderr_all_structures": float(need.sum() / rate),
        "mean_specularity_grad_median": float(np.median(grad)),
        "mean_specularity_grad_stderr_median": float(np.median(gse)),
        "gradient_components": nstruct,
        "paths_for_2sigma_gradient": n2s,
        "kappa_gradient_evals_per_s": rate / n2s if n2s else None,
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_bytes_per_path_profiled": traffic / paths_launch if traffic else None,
                     "algorithmic_bytes_per_path": alg_bytes, "bvh_nodes_per_path": nodes,
                     "leaf_tests_per_path": tests,
                     "valu_issue_fraction_profiled": prof.get("valu_issue_fraction"),
                     "l2_hit_rate_profiled": prof.get("l2_hit_rate")},
        "autotune": {k: v for k, v in tune.items() if k != "table"},
        "autotune_configs": len(tune["table"]),
        "host_generate_s": t_gen, "host_build_s": t_build,
        "cpu_baseline": None,
    }
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
            O.trace_alias(descs[s][0], mdesc, t.mfp_nm, prob, alias, 5000, first, 500)
            return 500
        out["cpu_baseline"] = cpu_leg_baseline(work, units, cpu, "16 structures (every 16th) x "
                                               "up to 1e4 paths each, brute-force candidate scans, "
                                               "FP64, 500 per task")
        out["vs_cpu"] = rate / out["cpu_baseline"]["value"]
    return out


# ------------------------------------------------------------------------------ main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--paths", type=float, default=1e9, help="paths per headline step")
    ap.add_argument("--count-paths", type=float, default=2e7,
                    help="paths of the op-counting FP64 launch of the headline workload")
    ap.add_argument("--compare-paths", type=float, default=2e8,
                    help="paths traced in all three precisions for the kappa comparison "
                         "(0: none, for profiling runs)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="CPU baseline sample of the headline (the legs take half)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every CPU of the host")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--film-np", type=int, default=834, help="film particles per cell (1e8 paths)")
    ap.add_argument("--film-steps", type=int, default=0, help="0: --steps")
    ap.add_argument("--adjoint-paths", type=float, default=1e8, help="0: skip the adjoint leg")
    ap.add_argument("--batch-structures", type=int, default=256, help="0: skip the batch leg")
    ap.add_argument("--batch-paths", type=float, default=1e5, help="paths per structure per launch")
    ap.add_argument("--batch-replays", type=int, default=1,
                    help="launches of the batch (fresh keys): 100 x 1e5 = the configured 1e7")
    ap.add_argument("--batch-graph", action="store_true",
                    help="replay the batch as a HIP graph (its tail kernel after the persistent one)")
    ap.add_argument("--headline-only", action="store_true", help="skip every leg")
    ap.add_argument("--only", choices=("film", "adjoint", "batch"), default=None,
                    help="run one leg alone, without the headline (profiling runs)")
    ap.add_argument("--batch-leaf", type=int, default=0,
                    help="BVH leaf size of the batch leg (0: the autotuner's choice)")
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

    import torch  # noqa: F401  (first: libptrans then shares torch's HIP runtime)
    from ptrans.trace import Context
    hc = host_cpu()
    threads = args.cpu_threads or hc["usable_cpus"]
    cpu_head = None if args.no_cpu else (args.cpu_seconds, threads)
    cpu_leg = None if args.no_cpu else (max(1.0, args.cpu_seconds / 2), threads)
    npaths = int(args.paths)
    with Context(seed=2024, device=0) as ctx:
        if args.only:
            if args.only == "film":
                res = film_leg(ctx, args.film_np, args.film_steps or args.steps, args.warmup, cpu_leg)
            elif args.only == "adjoint":
                res = adjoint_leg(ctx, int(args.adjoint_paths), cpu_leg)
            else:
                res = batch_leg(ctx, args.batch_structures, int(args.batch_paths),
                                args.batch_replays, cpu_leg, leaf=args.batch_leaf,
                                graph=args.batch_graph)
            print(json.dumps({"leg": args.only, "result": res}))
            return
        head = headline(ctx, npaths, args.steps, args.warmup, int(args.count_paths),
                        int(args.compare_paths), cpu_head)
        film = adj = bat = None
        if not args.headline_only:
            film = film_leg(ctx, args.film_np, args.film_steps or args.steps, args.warmup, cpu_leg)
            if args.adjoint_paths > 0:
                adj = adjoint_leg(ctx, int(args.adjoint_paths), cpu_leg)
                say("adjoint: %.1f ms" % adj["adjoint_kernel_ms"])
            if args.batch_structures > 0:
                bat = batch_leg(ctx, args.batch_structures, int(args.batch_paths),
                                args.batch_replays, cpu_leg, leaf=args.batch_leaf,
                                graph=args.batch_graph)
    dt = head.pop("dt")
    value = npaths * args.steps / dt
    cpu = head.pop("cpu_baseline")
    if cpu:
        cpu["host"] = hc
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
        "dtype": "f64",
        "data": "synthetic (seeded Voronoi geometry and bulk table)",
        "config": {"workload": head.pop("workload"), "paths_per_step": npaths,
                   "precision": "PT_PREC_FP32_FILTER: FP32 intersection bounds certified in FP64; "
                                "every result computed in FP64, bit-identical to the FP64 trace",
                   "launch": "pt_trace_alias_grains: first-pass pre-pass, persistent trace kernel, "
                             "tail kernel with 16 live workgroups beside it (long random walks "
                             "handed over after 500 events) and after it; defaults of "
                             "pt_ctx_set_tail and pt_ctx_set_tuning"},
        "roofline": head.pop("roofline"),
        "cpu_baseline": cpu,
        "speedup_vs_cpu_baseline": value / cpu["value"] if cpu else None,
        "kappa_gradient_evals_per_s": adj["kappa_gradient_evals_per_s"] if adj else None,
    }
    out.update(head)
    out["legs"] = {"film": film, "adjoint": adj, "batch": bat}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
