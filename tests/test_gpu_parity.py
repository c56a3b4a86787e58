"""libptrans on the GPU against the oracle and the reference's golden vectors.

Per path (pt_replay): exit side, event count, draw count, final position and direction
are bit-identical to the oracle with the shared transcendentals (the GPU uses the same
pt_math.h), and exit/draws/events/REAL(4) position equal what the reference wrote.
Per cell (pt_trace_grid): counts equal the reference's TransmissionFunction counts.
At larger sizes: closed forms and invariants.
"""
import numpy as np
import pytest

from _golden import CASES, Case
from oracle import oracle as O
from ptrans import gen, geometry
from ptrans.anphon import synthetic_table

pytestmark = pytest.mark.gpu


def _gpu_case(ctx, c, seed):
    ctx.set_seed(seed)
    g = ctx.geometry(c.geom)
    t = c.table
    m = ctx.modes(omega=t.omega, max_omega=t.max_omega)
    return g, m


@pytest.mark.parametrize("name", CASES)
def test_replay_bitexact_vs_oracle_and_reference(gpu_ctx, name):
    c = Case(name)
    mt = c.meta
    print("case", name, flush=True)
    g, m = _gpu_case(gpu_ctx, c, mt["paths_seed"])
    rec = gpu_ctx.replay(g, m, mt["mfp"], mt["nangle"], mt["nphonon"], 0, mt["npaths"])
    orc = O.replay(c.gdesc, c.mdesc, mt["mfp"], mt["nangle"], mt["nphonon"], mt["paths_seed"],
                   0, mt["npaths"], math=O.MATH_OWN)
    for f in ("end_side", "flags", "events", "draws", "which_grain", "start_grain"):
        np.testing.assert_array_equal(rec[f], orc[f], err_msg=f)
    np.testing.assert_array_equal(rec["pos"].view(np.uint64), orc["pos"].view(np.uint64))
    np.testing.assert_array_equal(rec["dir"].view(np.uint64), orc["dir"].view(np.uint64))
    ref = c.ref("own")
    np.testing.assert_array_equal((rec["end_side"] != 1).astype(np.int8), ref["trans"])
    np.testing.assert_array_equal(rec["draws"], ref["draws"])
    np.testing.assert_array_equal(rec["events"], np.diff(ref["traj_off"]) - 1)
    last = ref["traj"][ref["traj_off"][1:] - 1]
    np.testing.assert_array_equal(rec["pos"].astype(np.float32), last)


# Largest |REAL(4) position difference| (nm) against the reference run with the C
# library's transcendentals.  The GPU (like the oracle) uses its own log and sincos, which
# differ from glibc's by an ulp now and then.  On every case but one that shows only as
# float32 ulps of coordinates next to a wall (~1e-13 nm).  On the diffusive wire
# (c1_wire_gray: Lambda = 200 nm, p = 0 walls, up to ~350 events per path) the diffuse
# reflections amplify a 1e-14 nm difference by ~1.08 per event, to 0.016 nm after 349
# events (path 60) -- 1.6e-5 of the box, with every exit side, draw and event unchanged.
LIBM_POS_TOL = {"c1_wire_gray": 0.02}


@pytest.mark.parametrize("name", CASES)
def test_replay_vs_reference_with_libm_transcendentals(gpu_ctx, name):
    c = Case(name)
    mt = c.meta
    print("case", name, flush=True)
    g, m = _gpu_case(gpu_ctx, c, mt["paths_seed"])
    rec = gpu_ctx.replay(g, m, mt["mfp"], mt["nangle"], mt["nphonon"], 0, mt["npaths"])
    ref = c.ref("libm")
    np.testing.assert_array_equal((rec["end_side"] != 1).astype(np.int8), ref["trans"])
    np.testing.assert_array_equal(rec["draws"], ref["draws"])
    np.testing.assert_array_equal(rec["events"], np.diff(ref["traj_off"]) - 1)
    last = ref["traj"][ref["traj_off"][1:] - 1].astype(np.float64)
    err = np.abs(rec["pos"].astype(np.float32).astype(np.float64) - last)
    assert err.max() <= LIBM_POS_TOL.get(name, 1.0e-12), (name, err.max())


@pytest.mark.parametrize("name", CASES)
def test_grid_counts_vs_reference(gpu_ctx, name):
    c = Case(name)
    mt = c.meta
    print("case", name, flush=True)
    g, m = _gpu_case(gpu_ctx, c, mt["grid_seed"])
    counts, stats = gpu_ctx.trace_grid(g, m, mt["mfp"], mt["nangle"], mt["nphonon"])
    np.testing.assert_array_equal(counts, c.ref("own")["grid"])
    assert stats["paths"] == mt["nbranch"] * mt["nkpoint"] * (mt["nangle"] - 1) * mt["nphonon"]
    assert stats["lost"] == 0 and stats["start_mismatch"] == 0


def test_grid_start_mismatch_exits_cell_like_oracle(gpu_ctx):
    """A Voronoi generator moved 10% toward another: some start points then pick an
    incident grain that does not contain them, and the reference EXITs the cell's particle
    loop at the first such path (RT:1673-1682).  The GPU traces every path and its
    reduction counts only the transmitted paths before each cell's first mismatch: counts
    and skipped paths equal the oracle's, in both bit-identical precisions."""
    model = gen.voronoi_cube(seed=4, kind="Value", value=0.5, p=0.3, nphonon=300, nangle=5)
    geo = geometry.build_geometry(model)
    g2, g5 = np.array(geo.grains[2].generatrix), np.array(geo.grains[5].generatrix)
    geo.grains[5].generatrix = tuple(g5 + 0.1 * (g2 - g5))
    om = np.linspace(1e12, 9e13, 6).reshape(2, 3)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(om)
    oc, ost = O.trace_grid(gdesc, mdesc, 200.0, 5, 300, 17, math=O.MATH_OWN)
    assert ost[4] > 0 and ost[5] > 0                 # mismatching cells, skipped paths
    gpu_ctx.set_seed(17)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om)
    for prec in (0, FAST):
        counts, st = gpu_ctx.trace_grid(g, m, 200.0, 5, 300, precision=prec)
        print("precision", prec, "transmitted", counts.sum(), "skipped", st["skipped"], flush=True)
        np.testing.assert_array_equal(counts, oc)
        assert st["skipped"] == ost[5]


def test_grid_larger_vs_oracle_and_stats(gpu_ctx):
    """A few hundred thousand paths of the film-with-holes input with internal
    scattering: counts and the total event and draw counts equal the oracle's."""
    model = gen.film_holes(seed=2, nphonon=2000, nangle=7, mfp=(300.0,))
    geo = geometry.build_geometry(model)
    t = synthetic_table(3, 10, seed=5)
    om = t.f_thz / 0.03 * 0.03 * (3.14159265358979323846 * 2.0) * 1.0e12
    gdesc, _ga = geo.desc()
    mdesc, _ma = geometry.modes_desc(om, om, om, om)
    gpu_ctx.set_seed(99)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om)
    counts, stats = gpu_ctx.trace_grid(g, m, 300.0, 7, 2000)
    oc, ostats = O.trace_grid(gdesc, mdesc, 300.0, 7, 2000, 99, math=O.MATH_OWN)
    np.testing.assert_array_equal(counts, oc)
    assert stats["paths"] == ostats[0] and stats["events"] == ostats[1]
    assert stats["draws"] == ostats[2]


def test_specular_wire_transmits_everything(gpu_ctx):
    """Closed form: a ballistic wire with mirror walls (p = 1) transmits every phonon
    that enters, at every angle, so every traced cell counts exactly np."""
    model = gen.nanowire(p_wall=1.0, nphonon=5000, nangle=11)
    geo = geometry.build_geometry(model)
    om = np.full((2, 4), 1.0e13)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om)
    counts, stats = gpu_ctx.trace_grid(g, m, 1.0e10, 11, 5000)
    assert np.all(counts[:, :, :-1] == 5000)
    assert np.all(counts[:, :, -1] == 0)


def test_results_independent_of_run(gpu_ctx):
    """Counter-based streams: two runs of the same input give identical counts, and a
    different seed gives different counts."""
    model = gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.4, nphonon=500, nangle=5,
                           mfp=(500.0,))
    geo = geometry.build_geometry(model)
    om = np.linspace(1e12, 9e13, 8).reshape(2, 4)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om)
    gpu_ctx.set_seed(5)
    a, _ = gpu_ctx.trace_grid(g, m, 500.0, 5, 500)
    b, _ = gpu_ctx.trace_grid(g, m, 500.0, 5, 500)
    gpu_ctx.set_seed(6)
    c2, _ = gpu_ctx.trace_grid(g, m, 500.0, 5, 500)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c2)


# ---------------------------------------------------------------- FP32 filter + FP64 retest

FAST = 2       # PT_PREC_FP32_FILTER: FP32 intersection filter, every other step FP64
FP32 = 1       # PT_PREC_FP32_RETEST: FP32 filter and FP32 direction sampling/reflection


@pytest.mark.parametrize("name", CASES)
def test_fast_replay_identical_to_fp64(gpu_ctx, name):
    """The FP32-filter path certifies each event's surface (or re-tests it in FP64) and
    recomputes its distance in FP64, so every path evolves exactly as in FP64."""
    c = Case(name)
    mt = c.meta
    print("case", name, flush=True)
    g, m = _gpu_case(gpu_ctx, c, mt["paths_seed"])
    fast = gpu_ctx.replay(g, m, mt["mfp"], mt["nangle"], mt["nphonon"], 0, mt["npaths"], precision=FAST)
    ref = gpu_ctx.replay(g, m, mt["mfp"], mt["nangle"], mt["nphonon"], 0, mt["npaths"])
    for f in ("end_side", "flags", "events", "draws", "which_grain"):
        np.testing.assert_array_equal(fast[f], ref[f], err_msg=f)
    np.testing.assert_array_equal(fast["pos"].view(np.uint64), ref["pos"].view(np.uint64))
    np.testing.assert_array_equal(fast["dir"].view(np.uint64), ref["dir"].view(np.uint64))


@pytest.mark.parametrize("builder", ["film", "inclusions", "voronoi"])
def test_fast_grid_counts_identical_to_fp64_large(gpu_ctx, builder):
    """Half a million to a few million paths: FP32-filter counts, events and draws equal the
    FP64 trace's exactly; the fraction of events needing the FP64 full scan is small.
    Voronoi cells are all polygons, which the filter bounds by their FP64 plane hits and
    tests for containment only when nearest."""
    if builder == "film":
        model = gen.film_holes(seed=2, nphonon=20000, nangle=11, mfp=(300.0,))
    elif builder == "inclusions":
        model = gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.4, nphonon=20000, nangle=11,
                               mfp=(500.0,))
    else:
        model = gen.voronoi_cube(seed=4, kind="Gamma", value=1.5, p=0.2, nphonon=20000, nangle=11,
                                 mfp=(150.0,))
    geo = geometry.build_geometry(model)
    om = np.linspace(1e12, 9e13, 6).reshape(2, 3)
    gpu_ctx.set_seed(31)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om)
    a, sa = gpu_ctx.trace_grid(g, m, model.mfp[0], 11, 20000)
    b, sb = gpu_ctx.trace_grid(g, m, model.mfp[0], 11, 20000, precision=FAST)
    print(builder, "events", sa["events"], "retests", sb["retests"], "fallbacks", sb["fallbacks"],
          "fp64 ms", sa["kernel_ms"], "fast ms", sb["kernel_ms"], flush=True)
    np.testing.assert_array_equal(a, b)
    assert sa["events"] == sb["events"] and sa["draws"] == sb["draws"]
    assert sb["fallbacks"] < 0.01 * sb["events"]


@pytest.mark.parametrize("builder", ["film", "inclusions", "voronoi"])
def test_fp32_trace_statistically_equals_fp64(gpu_ctx, builder):
    """PT_PREC_FP32_RETEST reflects and samples directions in FP32, so paths drift from
    the FP64 ones by rounding; with common random numbers more than 99% of paths still
    end on the same side, and the transmitted total agrees within 3 sigma of the paired
    difference and within 0.5%."""
    if builder == "film":
        model = gen.film_holes(seed=2, nphonon=20000, nangle=11, mfp=(300.0,))
    elif builder == "inclusions":
        model = gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.4, nphonon=20000, nangle=11,
                               mfp=(500.0,))
    else:
        model = gen.voronoi_cube(seed=4, kind="Gamma", value=1.5, p=0.2, nphonon=20000, nangle=11)
    geo = geometry.build_geometry(model)
    om = np.linspace(1e12, 9e13, 6).reshape(2, 3)
    mfp = model.mfp[0] if model.mfp else 1.0e10
    gpu_ctx.set_seed(37)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om)
    # paired sample: every 10th path of the whole grid (all 60 cells), both precisions
    _, a = gpu_ctx.trace_log(g, m, mfp, 11, 20000, 0, 120000, stride=10, max_points=1)
    _, b = gpu_ctx.trace_log(g, m, mfp, 11, 20000, 0, 120000, stride=10, max_points=1,
                             precision=FP32)
    ta = (a["end_side"] != 1)
    tb = (b["end_side"] != 1)
    agree = (a["end_side"] == b["end_side"]).mean()
    print(builder, "per-path exit agreement", agree, flush=True)
    assert agree > 0.99
    ca, sa = gpu_ctx.trace_grid(g, m, mfp, 11, 20000)
    cb, sb = gpu_ctx.trace_grid(g, m, mfp, 11, 20000, precision=FP32)
    n = sa["paths"]
    diff = int(cb.sum()) - int(ca.sum())
    d = (tb.astype(int) - ta.astype(int))
    sig = np.sqrt(max(d.var(), 1.0 / len(d)) * n)
    print(builder, "transmitted fp64", ca.sum(), "fp32", cb.sum(), "diff", diff, "3 sigma", 3 * sig,
          "fp64 ms", sa["kernel_ms"], "fp32 ms", sb["kernel_ms"], flush=True)
    assert abs(diff) <= 3 * sig + 1
    assert abs(diff) / max(1, ca.sum()) < 0.005


def test_fp32_directions_unit_length_on_specular_spheres(gpu_ctx):
    """Specular spheres that never transmit (t = 0), ballistic: the sphere distance
    (RT:2985-3038) assumes |dir| = 1, so an FP32-made direction must be renormalised in
    FP64 or the hit point lands inside the sphere and the phonon is trapped there.  No
    FP32 path may end inside a sphere or at the event limit, |dir| = 1 to FP64
    rounding, and the mean event count matches the FP64 trace's."""
    model = gen.inclusions(seed=3, nphonon=100, nangle=11, ncyl=0, nsphere=8, kind="Value",
                           value=0.0, p=1.0)
    geo = geometry.build_geometry(model)
    sph = [(np.asarray(s.center), s.radius) for gr in geo.grains for s in gr.surfaces
           if s.kind == 2]
    om = np.linspace(1e12, 9e13, 6).reshape(2, 3)
    gpu_ctx.set_seed(41)
    gpu_ctx.set_event_limit(20000)
    try:
        g = gpu_ctx.geometry(geo)
        m = gpu_ctx.modes(omega=om)
        a = gpu_ctx.replay(g, m, 1.0e10, 11, 100, 0, 2000)
        b = gpu_ctx.replay(g, m, 1.0e10, 11, 100, 0, 2000, precision=FP32)
    finally:
        gpu_ctx.set_event_limit(100000000)
    print("mean events fp64", a["events"].mean(), "fp32", b["events"].mean(), flush=True)
    for rec in (a, b):
        assert not np.any(rec["flags"] & 4)
        for c, r in sph:
            assert np.all(np.linalg.norm(rec["pos"] - c, axis=1) > r - 1.0e-6)
    assert np.max(np.abs(np.linalg.norm(b["dir"], axis=1) - 1.0)) < 1.0e-14
    assert abs(b["events"].mean() / a["events"].mean() - 1.0) < 0.05


def test_safety_ball_diffusive_paths_bitexact(gpu_ctx):
    """Deeply diffusive paths (2 nm mean free path in a 200 x 100 x 100 nm wire): the
    accelerated precisions skip the collision scan while a free flight stays inside the
    path's safety ball; per-path records still equal the oracle's (which always scans) bit
    for bit, and the counts of a larger run equal the FP64 trace's."""
    model = gen.nanowire(length=200.0, width=100.0, height=100.0, p_wall=0.4, nphonon=200,
                         nangle=5)
    geo = geometry.build_geometry(model)
    om = np.linspace(1e12, 9e13, 2).reshape(2, 1)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(om)
    gpu_ctx.set_seed(23)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om)
    gpu_ctx.set_event_limit(10_000_000)
    try:
        rec = gpu_ctx.replay(g, m, 2.0, 5, 200, 0, 1200, precision=FAST)
        orec = O.replay(gdesc, mdesc, 2.0, 5, 200, 23, 0, 1200, math=O.MATH_OWN)
        print("mean events", rec["events"].mean(), "max", rec["events"].max(), flush=True)
        assert rec["events"].mean() > 20
        for f in ("end_side", "flags", "events", "draws", "which_grain"):
            np.testing.assert_array_equal(rec[f], orec[f], err_msg=f)
        np.testing.assert_array_equal(rec["pos"].view(np.uint64), orec["pos"].view(np.uint64))
        np.testing.assert_array_equal(rec["dir"].view(np.uint64), orec["dir"].view(np.uint64))
        a, sa = gpu_ctx.trace_grid(g, m, 2.0, 5, 20000)
        b, sb = gpu_ctx.trace_grid(g, m, 2.0, 5, 20000, precision=FAST)
        print("events", sa["events"], "fp64 ms", sa["kernel_ms"], "filter+ball ms", sb["kernel_ms"],
              flush=True)
        np.testing.assert_array_equal(a, b)
        assert sa["events"] == sb["events"] and sa["draws"] == sb["draws"]
    finally:
        gpu_ctx.set_event_limit(100_000_000)


def test_diffusive_grid_fp32_trace_statistically_equals_fp64(gpu_ctx):
    """A grid trace with a 3 nm sampling mean free path in a 1000 x 100 x 100 nm wire runs
    the kernel variant that serves the last paths of each wave with the whole wave (safety
    balls form below 5 nm).  Paths of ~1e3 events decorrelate between the FP32 and FP64
    traces, so their transmitted totals agree within 3 sigma of the difference of two
    binomial counts; the bit-identical filter precision gives the same counts, events and
    draws as FP64."""
    model = gen.nanowire(length=1000.0, width=100.0, height=100.0, p_wall=0.0, nphonon=100,
                         nangle=11)
    geo = geometry.build_geometry(model)
    om = np.linspace(1e12, 9e13, 6).reshape(2, 3)
    gpu_ctx.set_seed(43)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om)
    a, sa = gpu_ctx.trace_grid(g, m, 3.0, 11, 1000)
    b, sb = gpu_ctx.trace_grid(g, m, 3.0, 11, 1000, precision=FAST)
    c, sc = gpu_ctx.trace_grid(g, m, 3.0, 11, 1000, precision=FP32)
    print("events", sa["events"], "transmitted fp64", a.sum(), "fp32", c.sum(), "ms", sa["kernel_ms"],
          sb["kernel_ms"], sc["kernel_ms"], flush=True)
    np.testing.assert_array_equal(a, b)
    assert sa["events"] == sb["events"] and sa["draws"] == sb["draws"]
    n, ta, tc = sa["paths"], int(a.sum()), int(c.sum())
    assert ta > 100
    sig = np.sqrt(n * (ta / n) * (1 - ta / n))
    assert abs(tc - ta) <= 3 * np.sqrt(2) * sig


def test_safety_ball_alias_inclusions_bitexact(gpu_ctx):
    """Alias mode on the C3 inclusions with mean free paths 1 - 40 nm: curved candidates
    bound the ball too; FP32-filter records equal the oracle's."""
    from test_alias import _weights
    model = gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.4, nphonon=1, nangle=3)
    geo = geometry.build_geometry(model)
    t = synthetic_table(2, 4, 9, mfp_min_nm=1.0, mfp_max_nm=40.0)
    omega = t.omega_cm() * 0.03 * 3.14159265358979323846 * 2.0 * 1.0e12
    gpu_ctx.set_seed(31)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    rec = gpu_ctx.replay_alias(g, m, 0, 600, precision=FAST)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(omega, t.vel, t.mfp_nm, t.weight)
    prob, alias = O.build_alias(_weights(omega, t.vel, t.mfp_nm, t.weight))
    _, _, orec = O.trace_alias(gdesc, mdesc, t.mfp_nm, prob, alias, 31, 0, 600, records=True)
    print("mean events", rec["events"].mean(), "max", rec["events"].max(), flush=True)
    for f in ("end_side", "flags", "events", "draws", "which_grain"):
        np.testing.assert_array_equal(rec[f], orec[f], err_msg=f)
    np.testing.assert_array_equal(rec["pos"].view(np.uint64), orec["pos"].view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("mfp", [3.0, 30.0])
def test_lone_path_in_ball_fast_path_bitexact(gpu_ctx, mfp):
    """Diffusive paths (one mode, mfp 3 or 30 nm, inclusions) traced by the persistent
    kernel, whose waves end with single paths running their in-ball events with the whole
    wave's help, against paths traced one per lane without it: identical per-mode tallies
    and identical event and draw totals; and, in the bit-exact FP32 filter, the oracle's."""
    from test_alias import _weights
    model = gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.5, nphonon=1, nangle=3)
    geo = geometry.build_geometry(model)
    om = np.array([[2e13]])
    vel, w = np.array([[5000.0]]), np.array([[1.0]])
    lam = np.array([[mfp]])
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om, vel=vel, mfp=lam, weight=w)
    n = 512
    for prec in (1, 2):
        gpu_ctx.set_seed(71)
        mp, mt, st = gpu_ctx.trace_alias(g, m, n, precision=prec)
        rec = gpu_ctx.replay_alias(g, m, 0, n, precision=prec)
        print("mfp", mfp, "prec", prec, "events", st["events"], flush=True)
        assert st["events"] == int(rec["events"].sum()) and st["draws"] == int(rec["draws"].sum())
        assert int(mt.sum()) == int((rec["end_side"] != 1).sum())
        if prec == 2:
            gdesc, _g = geo.desc()
            mdesc, _m = geometry.modes_desc(om, vel, lam, w)
            prob, alias = O.build_alias(_weights(om, vel, lam, w))
            op, ot, orec = O.trace_alias(gdesc, mdesc, lam, prob, alias, 71, 0, n, records=True)
            np.testing.assert_array_equal(mt, ot)
            assert st["events"] == int(orec["events"].sum())
            np.testing.assert_array_equal(rec["pos"].view(np.uint64), orec["pos"].view(np.uint64))
