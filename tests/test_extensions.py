"""Extensions of the configurations that the reference cannot express: the Lorentzian
boundary law t = 1/(1 + (omega/omega0)^2) (C4) and periodic side walls (C2).

Pinned on the CPU by limits and invariants of the oracle (omega0 -> infinity gives the
Value law t = 1 path for path, omega0 -> 0 gives t = 0; a y-invariant film transmits the
same with periodic and with mirror side walls), and on the GPU by per-path bit-equality
with the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptrans import gen, geometry
from ptrans.constants import TRANS_LORENTZ, TRANS_VALUE

OM = np.linspace(2e12, 9e13, 6).reshape(2, 3)


def _voronoi(law, param):
    m = gen.voronoi_cube(seed=4, kind="Value", value=0.5, p=0.3, nphonon=50, nangle=5)
    geo = geometry.build_geometry(m)
    for gi, g in enumerate(geo.grains):
        if gi == 0:
            continue
        for s in g.surfaces:
            if s.internal and s.grain_outside > 0:
                s.trans_kind, s.trans_param = law, param
    return geo


def _replay(geo, n=400, mfp=300.0, seed=5):
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(OM)
    return O.replay(gdesc, mdesc, mfp, 5, 50, seed, 0, n, math=O.MATH_OWN)


@pytest.mark.parametrize("omega0, value", [(1.0e300, 1.0), (1.0e-300, 0.0)])
def test_lorentz_limits_equal_value_law(omega0, value):
    a = _replay(_voronoi(TRANS_LORENTZ, omega0))
    b = _replay(_voronoi(TRANS_VALUE, value))
    for f in ("end_side", "events", "draws", "which_grain"):
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)
    np.testing.assert_array_equal(a["pos"], b["pos"])


def test_lorentz_intermediate_differs_and_is_between():
    """omega0 within the band: the transmitted total lies between the t = 0 and t = 1
    limits (a boundary that transmits some phonons)."""
    gdesc = {}
    tot = {}
    for name, geo in (("t0", _voronoi(TRANS_VALUE, 0.0)), ("t1", _voronoi(TRANS_VALUE, 1.0)),
                      ("lor", _voronoi(TRANS_LORENTZ, 2 * np.pi * 5e12))):
        gd, _g = geo.desc()
        md, _m = geometry.modes_desc(OM)
        c, _ = O.trace_grid(gd, md, 300.0, 5, 400, 9, math=O.MATH_OWN)
        tot[name] = int(c.sum())
    assert min(tot["t0"], tot["t1"]) < tot["lor"] < max(tot["t0"], tot["t1"]), tot


def _slab(law, param, L=1000.0, a=400.0, p=1.0, nphonon=100):
    """Two grains split at x = a, mirror side walls: a ballistic path meets the boundary
    once from grain 1 and leaves through x = L if it crosses, through x = 0 if not."""
    geo = geometry.build_geometry(gen.two_grain_slab(L=L, a=a, p=p, nphonon=nphonon))
    for g in geo.grains[1:]:
        for s in g.surfaces:
            if s.internal and s.grain_outside > 0:
                s.trans_kind, s.trans_param = law, param
    return geo


def test_lorentz_law_closed_form_on_a_slab():
    """Ballistic two-grain slab with a specular (p = 1) Lorentzian boundary: every path
    meets the boundary exactly once, so T(omega) = t(omega) = 1/(1 + (omega/omega0)^2)
    exactly in expectation.  Modes at omega0/2, omega0 and 2 omega0 (t = 0.8, 0.5, 0.2),
    both traced angles pooled: T is within 3 binomial sigma of t, and at least 8 sigma
    from the law without the square (0.667, 0.5, 0.333), so a dropped square fails."""
    w0 = 2 * np.pi * 5.0e12
    om = np.array([[0.5 * w0, w0, 2.0 * w0]])                  # nb = 1, nk = 3
    geo = _slab(TRANS_LORENTZ, w0)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(om)
    np_ = 20000
    counts, st = O.trace_grid(gdesc, mdesc, 1.0e10, 3, np_, 31, math=O.MATH_OWN)
    assert st[3] == 0 and st[4] == 0                           # none lost, no start mismatch
    n = 2 * np_
    for k, ratio in enumerate((0.5, 1.0, 2.0)):
        T = counts[0, k, :2].sum() / n
        t = 1.0 / (1.0 + ratio * ratio)
        sig = np.sqrt(t * (1 - t) / n)
        assert abs(T - t) <= 3 * sig, (ratio, T, t)
        wrong = 1.0 / (1.0 + ratio)
        if ratio != 1.0:
            assert abs(T - wrong) >= 8 * sig, (ratio, T, wrong)


def test_grain_flux_attribution_closed_form_on_a_slab():
    """Alias mode on the slab with a boundary of t = 1, one ballistic mode: every path
    transmits, crossing grain 1 (x in [0, a]) and then grain 2 ([a, L]) with |d_x| =
    cos(theta) kept by the mirror side walls.  Per path the flights in grain 1 sum to
    a in x and a / cos(theta) in length, those in grain 2 to L - a and (L - a) /
    cos(theta), up to the 1e-7 nm nudges (at most 1e-7 nm per event), with cos(theta) =
    sqrt(u) from the path's Lambertian draw (draw 4: two mode draws, y, z, then u).
    Crediting a flight to the grain after a crossing, or dropping the box-grain order,
    moves a, L - a or the lengths between the grains and fails."""
    L, a, lam = 1000.0, 400.0, 1.0e10
    geo = _slab(TRANS_VALUE, 1.0, L=L, a=a)
    om = np.array([[3.0e13]])
    mfp = np.array([[lam]])
    one = np.ones((1, 1))
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(om, one, mfp, one)
    prob, alias = O.build_alias(np.ones(1))
    n, seed = 4000, 23
    mp, mt, gt = O.trace_alias_grains(gdesc, mdesc, mfp, prob, alias, seed, 0, n, len(geo.grains))
    _, _, rec = O.trace_alias(gdesc, mdesc, mfp, prob, alias, seed, 0, n, records=True)
    assert int(mp.sum()) == n and int(mt.sum()) == n           # t = 1: all transmit
    assert np.all(rec["end_side"] == 2)
    inv_cos = np.array([1.0 / np.sqrt(O.philox_draw(seed, i, 4)) for i in range(n)])
    tol = 1.0e-7 * float(rec["events"].sum()) + 1e-9 * n * L
    dx, dl = gt[:, 0] * lam, gt[:, 1] * lam
    assert dx[0] == 0.0 and dl[0] == 0.0                       # never a flight in the box grain
    assert abs(dx[1] - n * a) <= tol and abs(dx[2] - n * (L - a)) <= tol, (dx, n * a)
    assert abs(dl[1] - a * inv_cos.sum()) <= tol, (dl[1], a * inv_cos.sum())
    assert abs(dl[2] - (L - a) * inv_cos.sum()) <= tol, (dl[2], (L - a) * inv_cos.sum())


def _film(periodic, holes):
    m = gen.film_holes(seed=2, nholes=holes, p_side=1.0, p_film=0.2, nphonon=4000, nangle=5)
    geo = geometry.build_geometry(m)
    if periodic:
        geometry.set_periodic(geo, axes=(1,))
    return geo


def test_periodic_equals_mirror_for_y_invariant_film():
    """Without holes nothing depends on y, so unfolding mirror (p = 1) side walls into a
    periodic cell changes no path's x, z motion: the transmitted totals agree within
    3 sigma (the two use different numbers of draws, so they are independent samples)."""
    res = []
    for periodic in (False, True):
        gd, _g = _film(periodic, 0).desc()
        md, _m = geometry.modes_desc(OM)
        c, st = O.trace_grid(gd, md, 300.0, 5, 4000, 11, math=O.MATH_OWN)
        res.append((int(c.sum()), int(st[0])))
    (t0, n), (t1, _) = res
    p = (t0 + t1) / (2.0 * n)
    sig = np.sqrt(2 * n * p * (1 - p))
    assert abs(t0 - t1) <= 3 * sig + 1, (t0, t1, sig)


def test_periodic_paths_stay_in_the_box():
    geo = _film(True, 3)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(OM)
    rec = O.replay(gdesc, mdesc, 300.0, 5, 4000, 3, 0, 2000, math=O.MATH_OWN)
    assert not np.any(rec["flags"] & 1)
    assert np.all(rec["pos"][:, 1] >= -1e-5) and np.all(rec["pos"][:, 1] <= geo.box[1] + 1e-5)


@pytest.mark.gpu
def test_periodic_geometry_validation(gpu_ctx):
    from ptrans import _lib
    geo = _film(False, 3)
    geo.grains[0].surfaces[2].trans_kind = 5          # one wall of the pair only
    with pytest.raises(_lib.PtransError, match="pairs"):
        gpu_ctx.geometry(geo)
    geo = _film(False, 3)
    geo.grains[0].surfaces[0].trans_kind = 5          # an exit wall
    with pytest.raises(_lib.PtransError, match="side walls"):
        gpu_ctx.geometry(geo)


# ------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("case", ["lorentz_voronoi", "periodic_film"])
def test_gpu_extension_paths_bitexact_vs_oracle(gpu_ctx, case):
    if case == "lorentz_voronoi":
        m = gen.voronoi_cube(seed=4, kind="Gamma", value=1.5, p=0.3, nphonon=50, nangle=5)
        geo = geometry.build_geometry(m)
        geometry.set_lorentz_boundaries(geo, np.random.default_rng(2))
        mfp = 400.0
    else:
        geo = _film(True, 3)
        mfp = 300.0
    print("case", case, flush=True)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(OM)
    orec = O.replay(gdesc, mdesc, mfp, 5, 50, 17, 0, 1000, math=O.MATH_OWN)
    gpu_ctx.set_seed(17)
    g = gpu_ctx.geometry(geo)
    mm = gpu_ctx.modes(omega=OM)
    for prec in (0, 2):
        rec = gpu_ctx.replay(g, mm, mfp, 5, 50, 0, 1000, precision=prec)
        for f in ("end_side", "flags", "events", "draws", "which_grain"):
            np.testing.assert_array_equal(rec[f], orec[f], err_msg="%s prec %d" % (f, prec))
        np.testing.assert_array_equal(rec["pos"].view(np.uint64), orec["pos"].view(np.uint64))
    oc, _ = O.trace_grid(gdesc, mdesc, mfp, 5, 50, 17, math=O.MATH_OWN)
    for prec in (0, 2):
        c, _ = gpu_ctx.trace_grid(g, mm, mfp, 5, 50, precision=prec)
        np.testing.assert_array_equal(c, oc)


def _c4(seed=2):
    m = gen.voronoi_cube(seed=4, kind="Gamma", value=1.5, p=0.3, nphonon=1, nangle=3)
    geo = geometry.build_geometry(m)
    geometry.set_lorentz_boundaries(geo, np.random.default_rng(seed))
    return geo


def _alias_inputs(nb=2, nk=5, seed=7):
    from test_alias import _table, _weights
    omega, vel, mfp, weight = _table(nb, nk, seed)
    mfp = np.clip(mfp, 50.0, None)
    prob, alias = O.build_alias(_weights(omega, vel, mfp, weight))
    return omega, vel, mfp, weight, prob, alias


def test_grain_flux_tallies_split_kappa():
    """Per-grain flux tallies add up to L * sum T / Lambda (a path's x displacements sum to
    L T up to the nudges), so the grain shares of kappa add up to kappa."""
    from ptrans import post
    geo = _c4()
    omega, vel, mfp, weight, prob, alias = _alias_inputs()
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(omega, vel, mfp, weight)
    mp, mt, gt = O.trace_alias_grains(gdesc, mdesc, mfp, prob, alias, 3, 0, 20000, len(geo.grains))
    total = geo.box[0] * float((mt / mfp).sum())
    assert abs(gt[:, 0].sum() - total) <= 1e-8 * total
    n = int(mp.sum())
    ratio = post.alias_kappa(mp, mt, mfp, geo.box[0]).ratio
    shares = post.grain_kappa_shares(gt, n)
    assert abs(shares.sum() - ratio) <= 1e-8 * ratio
    assert gt[0, 1] == 0.0 and np.all(gt[1:, 1] > 0)        # no flight in the box grain
    mp2, mt2 = O.trace_alias(gdesc, mdesc, mfp, prob, alias, 3, 0, 20000)
    np.testing.assert_array_equal(mp, mp2)
    np.testing.assert_array_equal(mt, mt2)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 2])
def test_gpu_grain_tallies_vs_oracle(gpu_ctx, precision):
    geo = _c4()
    omega, vel, mfp, weight, prob, alias = _alias_inputs()
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(omega, vel, mfp, weight)
    mp, mt, gt = O.trace_alias_grains(gdesc, mdesc, mfp, prob, alias, 3, 0, 20000, len(geo.grains))
    gpu_ctx.set_seed(3)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    print("grain tallies, precision", precision, flush=True)
    gp, gtr, gg, st = gpu_ctx.trace_alias_grains(g, m, 20000, precision=precision)
    np.testing.assert_array_equal(gp, mp)
    np.testing.assert_array_equal(gtr, mt)
    np.testing.assert_allclose(gg, gt, rtol=1e-11, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("periodic", [True, False])
def test_gpu_fp32_film_paths_all_end(gpu_ctx, periodic):
    """FP32 trace of the 600 x 600 x 200 nm film with 3 holes, 6 x 2000 modes x 10 angles x
    834 particles (1e8 paths): no path reaches a 1e6-event limit.  A diffuse reflection
    whose FP32 draw rounds to 1.0f leaves the surface exactly tangent; without the clamp
    of that draw, FP32 rounding of the frame sent such paths into a vacuum hole, where they
    bounced until the event limit (about one path in 1e8 on this film)."""
    from ptrans.anphon import synthetic_table
    from ptrans.constants import M_TWOPI
    m = gen.film_holes(seed=2, lx=600.0, ly=600.0, h=200.0, radius=50.0, nholes=3, p_hole=0.3,
                       p_film=0.0, p_side=1.0, nphonon=834, nangle=11)
    geo = geometry.build_geometry(m)
    if periodic:
        geometry.set_periodic(geo, axes=(1,))
    t = synthetic_table(6, 2000, seed=20)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    g = gpu_ctx.geometry(geo)
    mm = gpu_ctx.modes(omega=omega)
    gpu_ctx.set_seed(12345)
    gpu_ctx.set_event_limit(1_000_000)
    try:
        print("film periodic=%s, FP32 trace of 1e8 paths" % periodic, flush=True)
        counts, st = gpu_ctx.trace_grid(g, mm, 1.0e10, 11, 834, precision=1)
    finally:
        gpu_ctx.set_event_limit(100_000_000)
    print("events/path %.3f, kernel %.1f ms" % (st["events"] / st["paths"], st["kernel_ms"]), flush=True)
    assert st["paths"] == 6 * 2000 * 10 * 834
    assert st["event_limited"] == 0
    assert st["lost"] == 0
