"""Alias-mode estimator (an extension of the reference's (k, b, angle) enumeration):
the Vose table, the equivalence with the grid estimator for a gray table, and the GPU
against the oracle path by path and tally by tally."""
import numpy as np
import pytest

from _golden import Case
from oracle import oracle as O
from ptrans import gen, geometry, post, trace
from ptrans.anphon import synthetic_table

M_TWOPI = 3.14159265358979323846 * 2.0


def _table(nb, nk, seed, mfp=None):
    t = synthetic_table(nb, nk, seed)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    mfp_nm = t.mfp_nm if mfp is None else np.full((nb, nk), float(mfp))
    return omega, t.vel, mfp_nm, t.weight


def test_alias_table_library_equals_oracle_and_exact():
    rng = np.random.default_rng(4)
    for n in (1, 2, 7, 1000, 12000):
        w = rng.random(n) ** 4
        if n > 3:
            w[::5] = 0.0
        p1, a1 = trace.alias_table(w)
        p2, a2 = O.build_alias(w)
        np.testing.assert_array_equal(p1, p2)
        np.testing.assert_array_equal(a1, a2)
        P = trace.alias_probabilities(p1, a1)
        np.testing.assert_allclose(P, w / w.sum(), atol=1e-13)
        assert np.all(P[w == 0] == 0)


def test_alias_table_rejects_bad_weights():
    with pytest.raises(Exception):
        trace.alias_table([0.0, 0.0])
    with pytest.raises(Exception):
        trace.alias_table([1.0, -1.0])


def _weights(omega, vel, mfp, weight, T=300.0):
    cap = np.vectorize(lambda o: post.heat_capacity(float(o), T))(omega)
    return (cap * vel * mfp * weight).reshape(-1, order="F")


def test_sampled_mode_frequencies_follow_weights():
    """Chi-square of the modes drawn by 40000 oracle alias paths against C v Lambda w."""
    model = gen.nanowire(p_wall=1.0, nphonon=1, nangle=3)
    geo = geometry.build_geometry(model)
    omega, vel, mfp, weight = _table(2, 5, 3, mfp=1.0e10)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(omega, vel, mfp, weight)
    w = _weights(omega, vel, mfp, weight)
    prob, alias = O.build_alias(w)
    mp, mt = O.trace_alias(gdesc, mdesc, mfp, prob, alias, 5, 0, 40000)
    obs = mp.reshape(-1, order="F")
    exp = w / w.sum() * 40000
    chi2 = ((obs - exp) ** 2 / exp).sum()
    assert chi2 < 9 + 5 * np.sqrt(2 * 9)          # 9 degrees of freedom, ~5 sigma
    assert mt.sum() == mp.sum()                   # mirror walls, ballistic: all transmit


def test_alias_matches_grid_estimator_for_gray_table():
    """With every mode at the same Lambda, E[(3L/4) T / Lambda] must equal
    Lambda_eff/Lambda of the reference's estimator, Simpson over a fine angle grid
    (RT:2418-2451): within 3 combined standard errors plus the quadrature error."""
    lam = 150.0
    model = gen.nanowire(length=600.0, p_wall=0.2, nphonon=1, nangle=3)
    geo = geometry.build_geometry(model)
    omega, vel, mfp, weight = _table(1, 4, 9, mfp=lam)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(omega, vel, mfp, weight)
    w = _weights(omega, vel, mfp, weight)
    prob, alias = O.build_alias(w)
    mp, mt = O.trace_alias(gdesc, mdesc, mfp, prob, alias, 11, 0, 60000)
    a = post.alias_kappa(mp, mt, mfp, model.box[0])
    na, np_ = 41, 1500
    om1 = omega[:, :1]
    mdesc1, _m1 = geometry.modes_desc(om1)
    counts, _ = O.trace_grid(gdesc, mdesc1, lam, na, np_, 13)
    T = counts / float(np_)
    ang = post.vangle(na)
    lam_eff = post.effective_mfp(T, ang, model.box[0])[0, 0]
    # binomial error of the Simpson sum
    wts = np.zeros(na)
    for ia in range(1, na - 1, 2):
        wts[ia - 1] += 1; wts[ia] += 4; wts[ia + 1] += 1
    ds = (ang[1] - ang[0]) / 3.0 * 1.5 * model.box[0]
    sc = np.sin(ang) * np.cos(ang)
    var = ((T * (1 - T) / np_)[0, 0] * (wts * sc * ds) ** 2).sum()
    sig = np.sqrt(a.stderr ** 2 + var / lam ** 2)
    assert abs(a.ratio - lam_eff / lam) < 3 * sig + 0.004, (a.ratio, lam_eff / lam, sig)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_wire_gray", "c2_film_holes_gray", "c3_inclusions_gamma",
                                  "c4_voronoi_gamma", "c5_tri_brick"])
def test_gpu_alias_paths_bitexact_vs_oracle(gpu_ctx, name):
    c = Case(name)
    print("case", name, flush=True)
    omega = c.table.omega
    _, vel, mfp, weight = _table(c.table.nbranch, c.table.nkpoint, 17)
    scale = np.linspace(0.7, 1.3, c.table.nbranch)
    gpu_ctx.set_seed(77)
    g = gpu_ctx.geometry(c.geom)
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight, max_omega=c.table.max_omega)
    rec = gpu_ctx.replay_alias(g, m, 0, 400, branch_scale=scale)
    mdesc, _m = geometry.modes_desc(omega, vel, mfp, weight, c.table.max_omega)
    prob, alias = O.build_alias(_weights(omega, vel, mfp, weight))
    mp, mt, orec = O.trace_alias(c.gdesc, mdesc, mfp, prob, alias, 77, 0, 400, scale=scale,
                                 records=True)
    for f in ("end_side", "flags", "events", "draws", "which_grain", "start_grain"):
        np.testing.assert_array_equal(rec[f], orec[f], err_msg=f)
    np.testing.assert_array_equal(rec["pos"].view(np.uint64), orec["pos"].view(np.uint64))
    np.testing.assert_array_equal(rec["dir"].view(np.uint64), orec["dir"].view(np.uint64))
    gp, gt, st = gpu_ctx.trace_alias(g, m, 400, branch_scale=scale)
    np.testing.assert_array_equal(gp, mp)
    np.testing.assert_array_equal(gt, mt)
    fast = gpu_ctx.replay_alias(g, m, 0, 400, branch_scale=scale, precision=2)
    for f in ("end_side", "events", "draws"):
        np.testing.assert_array_equal(fast[f], orec[f], err_msg="fast " + f)
    np.testing.assert_array_equal(fast["pos"].view(np.uint64), orec["pos"].view(np.uint64))


@pytest.mark.gpu
def test_gpu_alias_tallies_vs_oracle_large(gpu_ctx):
    model = gen.film_holes(seed=2, nphonon=1, nangle=3)
    geo = geometry.build_geometry(model)
    omega, vel, mfp, weight = _table(3, 50, 21)
    gpu_ctx.set_seed(5)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    gp, gt, st = gpu_ctx.trace_alias(g, m, 200000)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(omega, vel, mfp, weight)
    prob, alias = O.build_alias(_weights(omega, vel, mfp, weight))
    mp, mt = O.trace_alias(gdesc, mdesc, mfp, prob, alias, 5, 0, 200000)
    np.testing.assert_array_equal(gp, mp)
    np.testing.assert_array_equal(gt, mt)
    assert st["paths"] == 200000
