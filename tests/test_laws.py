"""Surface-law set-up: GammaDeterm (TBC -> gamma, RT:2499-2599) and the
FluctuationSpecularity jitter (RT:1151-1157), against the reference's own loadMCP on the
same inputs (tests/golden/laws, written by tests/golden/make_golden.py through the replay
driver's 'laws' mode) and against what the mathematics fixes."""
import os

import numpy as np
import pytest

import _golden  # noqa: F401  (sys.path)
from ptrans import anphon, geometry, laws, mcp
from ptrans.constants import TRANS_TBC

LAWS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "laws")
CASES = sorted(os.listdir(LAWS))


def _ref(case, seed):
    t, s = {}, {}
    with open(os.path.join(LAWS, case, "ref_seed%d.txt" % seed)) as f:
        for ln in f:
            w = ln.split()
            if w[0] == "T":
                t[(int(w[1]), int(w[2]))] = (int(w[3]), float(w[4]))
            else:
                s[(int(w[1]), int(w[2]))] = float(w[3])
    return t, s


def _case(case):
    d = os.path.join(LAWS, case)
    m = mcp.read_mcp(os.path.join(d, "input.mcp"))
    tab = anphon.read_anphon(os.path.join(d, "Si.anphon"), "Si")
    seeds = sorted(int(f[8:-4]) for f in os.listdir(d) if f.startswith("ref_seed"))
    return m, tab, seeds


@pytest.mark.parametrize("case", CASES)
def test_gamma_determ_equals_reference(case):
    m, tab, seeds = _case(case)
    t, _ = _ref(case, seeds[0])
    row = m.trans_rows[0]
    g = laws.gamma_determ(tab, row.value * 1.0e6, m.temperature)
    kind, gref = t[(1, 1)]
    assert kind == TRANS_TBC
    assert g == gref, (g, gref)                       # bit-for-bit (same sums, same order)


@pytest.mark.parametrize("case", CASES)
def test_surface_specularities_equal_reference(case):
    """Every surface of every grain (box, grains, fill copies) after loadMCP, for every
    set-up seed the reference ran with."""
    m, tab, seeds = _case(case)
    for seed in seeds:
        _, sref = _ref(case, seed)
        geo = geometry.build_geometry(m, tables={"Si": tab}, setup_seed=seed)
        got = {(ig, i + 1): s.specularity for ig, g in enumerate(geo.grains)
               for i, s in enumerate(g.surfaces)}
        internal = {(ig, i + 1) for ig, g in enumerate(geo.grains) for i, s in enumerate(g.surfaces)
                    if s.internal or ig == 0}
        assert set(sref) == set(got)
        for k in internal:                            # non-internal facets: never set (RT:1130)
            assert got[k] == sref[k], (case, seed, k, got[k], sref[k])


def test_gamma_inverts_the_boundary_conductance():
    """G(gamma) = TBC to Newton's tolerance over four decades; G falls as gamma grows."""
    t = anphon.synthetic_table(3, 60, 7)
    tab = anphon.Anphon(material="Si", rtemp=t.rtemp, vol=t.vol,
                        omega=t.omega_cm() * 0.03 * 2.0 * np.pi * 1.0e12, tau=t.tau_ps,
                        vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    prev = None
    for tbc in (1.0, 10.0, 100.0, 1000.0, 1.0e4):
        g = laws.gamma_determ(tab, tbc * 1.0e6, 300.0)
        assert abs(laws.boundary_conductance(tab, g, 300.0) / (tbc * 1.0e6) - 1.0) < 2e-5
        if prev is not None:
            assert g < prev
        prev = g


def test_gamma_determ_planted_error_is_caught():
    """A dropped factor (the /2 of the conductance sum) moves gamma far from the
    reference's value, so the golden comparison above would fail."""
    m, tab, seeds = _case("voronoi_tbc300_fluct")
    t, _ = _ref("voronoi_tbc300_fluct", seeds[0])
    half = anphon.Anphon(material="x", rtemp=tab.rtemp, vol=tab.vol / 2, omega=tab.omega,
                         tau=tab.tau, vel=tab.vel, mfp=tab.mfp, weight=tab.weight)
    g = laws.gamma_determ(half, m.trans_rows[0].value * 1.0e6, m.temperature)
    assert abs(g - t[(1, 1)][1]) > 1e-3


def test_fluctuation_is_irwin_hall():
    """The jitter is (sum of 12 uniforms - 6) sigma: mean 0, variance sigma^2, |x| <= 6 sigma."""
    u = laws.setup_uniforms(5, 0, 12 * 20000).reshape(-1, 12)
    x = np.array([laws.fluctuate_specularity(0.5, 0.01, r) - 0.5 for r in u])
    assert abs(x.mean()) < 4 * 0.01 / np.sqrt(len(x))
    assert abs(x.std() / 0.01 - 1.0) < 0.03
    assert np.all(np.abs(x) <= 0.06 + 1e-15)
    assert laws.fluctuate_specularity(0.02, 1.0, np.zeros(12)) == 0.0       # clipped
    assert laws.fluctuate_specularity(0.98, 1.0, np.ones(12) * 0.99) == 1.0


def test_setup_stream_is_the_library_stream():
    """pt_philox_fill gives the draws of stream 2^62, equal to the oracle's Philox."""
    from oracle import oracle as O
    u = laws.setup_uniforms(9, 3, 5)
    assert [float(x) for x in u] == [O.philox_draw(9, 1 << 62, d) for d in range(3, 8)]


def test_tbc_and_fluctuation_need_their_inputs():
    m, tab, _ = _case("voronoi_tbc300_fluct")
    with pytest.raises(ValueError, match="bulk table"):
        geometry.build_geometry(m, setup_seed=1)
    with pytest.raises(ValueError, match="setup_seed"):
        geometry.build_geometry(m, tables={"Si": tab})
