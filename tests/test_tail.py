"""Long random walks and the last few paths of a wave.  In the accelerated precisions a
path past `tail_events` events (pt_ctx_set_tail), or one of the last four of a wave once the
work queue is empty, leaves the persistent kernel for the tail kernel, which runs it on a
whole wave: it sets up a pending safety ball, runs batches of 64 in-ball events (their draws
computed 64 at a time, the positions by a sequential FP64 chain through LDS) and spreads the
next candidate scan over the lanes.  Diffusive modes make such paths common, so these tests
trace them at the FP32-filter precision, whose paths are bit-identical to FP64, and pin
everything such a path updates against the oracle: adjoint scores (scattering counts and
path lengths), per-grain flux tallies (flights inside one grain) and the per-mode tallies.
They run with the default hand-over (500 events), with every path handed over after its
first event, with a tail buffer smaller than the paths that ask for it (the rest stay in the
persistent kernel) and with no hand-over at all; with the tail kernel's live workgroups
beside the persistent kernel (16 by default, 64 at most: records taken while they are
written, the two tail kernels sharing each part's next index) and without them.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptrans import gen, geometry
from ptrans.anphon import synthetic_table
from test_adjoint import OracleSetup, _weights

M_TWOPI = 3.14159265358979323846 * 2.0
FILTER = 2
# (events, max_records, live_blocks) of pt_ctx_set_tail: the default, every path after one
# event, a buffer too small for the paths that ask (4096 records), no tail kernel; the first
# two with no live tail kernel, and every path with the most live workgroups
TAILS = [(500, 1 << 20, 16), (1, 1 << 20, 16), (1, 4096, 16), (500, 0, 16), (500, 1 << 20, 0),
         (1, 1 << 20, 0), (1, 1 << 20, 64)]
TAIL_IDS = ["default", "all", "overflow", "none", "after", "all_after", "all_live64"]


@pytest.fixture
def tail_ctx(gpu_ctx, request):
    gpu_ctx.set_tail(*request.param)
    yield gpu_ctx
    gpu_ctx.set_tail(*TAILS[0])


@pytest.mark.gpu
@pytest.mark.parametrize("tail_ctx", TAILS, ids=TAIL_IDS, indirect=True)
def test_gpu_adjoint_diffusive_tail_equals_oracle(tail_ctx):
    """Mean free paths of 2 - 20 nm in a 300 x 100 x 100 nm wire (safety balls below 5 nm),
    branch scales 0.8 and 1.25: adjoint batch sums equal the oracle's."""
    model = gen.nanowire(length=300.0, width=100.0, height=100.0, p_wall=0.5, nphonon=1, nangle=3)
    nsurf = len(geometry.build_geometry(model).flatten()["surfaces"])
    group = [-1] * nsurf
    group[2] = group[3] = 0
    group[4] = group[5] = 1
    s = OracleSetup(model, 2, 3, 11, (2.0, 20.0), [0.35, 0.6], group=group)
    scale = np.array([0.8, 1.25])
    tail_ctx.set_seed(19)
    g = tail_ctx.geometry(s.geo)
    g.set_specularity(s.p0)
    m = tail_ctx.modes(omega=s.omega, vel=s.vel, mfp=s.mfp, weight=s.weight)
    mp, mt, gs, st = tail_ctx.adjoint_alias(g, m, 20000, surface_group=s.group, ngroup=s.ngroup,
                                           baseline=0.002, nbatch=16, branch_scale=scale,
                                           precision=FILTER)
    print("events per path", st["events"] / st["paths"], flush=True)
    assert st["events"] / st["paths"] > 30
    os_ = s.sums(19, 20000, 16, baseline=0.002, scale=scale)
    np.testing.assert_array_equal(gs[:, 0], os_[:, 0])
    np.testing.assert_allclose(gs, os_, rtol=1e-11, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("tail_ctx", TAILS, ids=TAIL_IDS, indirect=True)
def test_gpu_grain_tallies_diffusive_tail_equal_oracle(tail_ctx):
    """Mean free paths of 3 - 30 nm in a 300 nm Voronoi cube with Lorentzian boundaries
    (safety balls below 15 nm): per-mode and per-grain tallies equal the oracle's."""
    m_ = gen.voronoi_cube(seed=4, L=300.0, ngrain=5, kind="Gamma", value=1.5, p=0.3, nphonon=1,
                          nangle=3)
    geo = geometry.build_geometry(m_)
    geometry.set_lorentz_boundaries(geo, np.random.default_rng(3))
    t = synthetic_table(2, 5, 7, mfp_min_nm=3.0, mfp_max_nm=30.0)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    prob, alias = O.build_alias(_weights(omega, t.vel, t.mfp_nm, t.weight))
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(omega, t.vel, t.mfp_nm, t.weight)
    mp, mt, gt = O.trace_alias_grains(gdesc, mdesc, t.mfp_nm, prob, alias, 23, 0, 20000,
                                      len(geo.grains))
    tail_ctx.set_seed(23)
    g = tail_ctx.geometry(geo)
    m = tail_ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    gp, gtr, gg, st = tail_ctx.trace_alias_grains(g, m, 20000, precision=FILTER)
    print("events per path", st["events"] / st["paths"], flush=True)
    assert st["events"] / st["paths"] > 30
    np.testing.assert_array_equal(gp, mp)
    np.testing.assert_array_equal(gtr, mt)
    np.testing.assert_allclose(gg, gt, rtol=1e-11, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("tail_ctx", TAILS[:2] + TAILS[4:], ids=TAIL_IDS[:2] + TAIL_IDS[4:],
                         indirect=True)
def test_gpu_many_candidates_tail_scan_equals_oracle(tail_ctx):
    """80 spheres in a 1000 nm cube: a path in the matrix has 86 candidates, so the wave's
    scan for a path left alone takes two passes of 64 lanes (the first strict minimum kept
    across passes).  Per-mode tallies equal the oracle's."""
    m_ = gen.inclusions(seed=7, L=1000.0, nsphere=80, ncyl=0, rs=(20.0, 40.0), kind="Gamma",
                        value=1.0, p=0.5, nphonon=1, nangle=3)
    geo = geometry.build_geometry(m_)
    assert len(geo.grains[-1].surfaces) == 80
    t = synthetic_table(2, 5, 11, mfp_min_nm=20.0, mfp_max_nm=400.0)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    prob, alias = O.build_alias(_weights(omega, t.vel, t.mfp_nm, t.weight))
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(omega, t.vel, t.mfp_nm, t.weight)
    mp, mt = O.trace_alias(gdesc, mdesc, t.mfp_nm, prob, alias, 29, 0, 20000)
    tail_ctx.set_seed(29)
    g = tail_ctx.geometry(geo)
    m = tail_ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    gp, gt, st = tail_ctx.trace_alias(g, m, 20000, precision=FILTER)
    print("events per path", st["events"] / st["paths"], "fallbacks", st["fallbacks"], flush=True)
    np.testing.assert_array_equal(gp, mp)
    np.testing.assert_array_equal(gt, mt)


_BATCH_REF = {}      # oracle sums of the batch case, shared by its parameters


@pytest.mark.gpu
@pytest.mark.parametrize("tail_ctx", [TAILS[0], TAILS[1], TAILS[4]], ids=["default", "all", "after"],
                         indirect=True)
def test_gpu_batch_long_walks_equal_oracle(tail_ctx, request):
    """Mean free paths of 1.2 - 4 nm in two structures of the screening generator (1,920 and
    12,000 triangles; walks of up to 7e5 events): the paths below 2 nm start first (their
    own queue), and the long walks run in the tail kernel, beside the persistent kernel and
    after it, with local candidate lists, scans limited to the free path and wave-cooperative
    traversals.  Per-structure path counts and sums of T / mfp equal the oracle's
    brute-force trace of each structure alone, in the bit-identical FP32 filter; with the
    default hand-over also as a graph replay (no first pass, the tail kernel after the
    persistent one)."""
    from test_batch import _oracle_struct_sums
    from ptrans.trace import Batch
    geos = [gen.sc_lattice_flat(seed=s, k=4, m=m_, jitter=5.0, trans_kind=2, trans_param=2.0, p=0.3)
            for s, m_ in ((402, 2), (415, 5))]
    t = synthetic_table(2, 4, 9, mfp_min_nm=1.2, mfp_max_nm=4.0)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    prob, alias = O.build_alias(_weights(omega, t.vel, t.mfp_nm, t.weight))
    mods = (omega, t.vel, t.mfp_nm, t.weight, prob, alias)
    npaths = [4000, 1000]
    tail_ctx.set_seed(51)
    m = tail_ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
    b = Batch(tail_ctx, geos, npaths, leaf_size=2, grid=32)
    runs = [b.trace(m, precision=FILTER)]
    if request.node.callspec.id == "default":
        runs.append(b.trace(m, precision=FILTER, graph=True))
    st = runs[0][2]
    print("events per path", st["events"] / st["paths"], flush=True)
    assert st["events"] / st["paths"] > 100
    for s in range(2):
        if s not in _BATCH_REF:
            _BATCH_REF[s] = _oracle_struct_sums(geos[s], mods, 51, b.path_off[s], npaths[s])
        n, sx, _nt = _BATCH_REF[s]
        for sums, _fin, _st in runs:
            assert sums[s, 0] == n, (s, sums[s, 0], n)
            assert abs(sums[s, 1] - sx) <= 1e-12 * max(sx, 1e-300), (s, sums[s, 1], sx)
    b.close()
