"""The event safety limit (no counterpart in the reference, whose event loop has no bound,
RT:1693): a path stopped by it has not ended on either side, so every tally counts it as
traced and not transmitted, and the stats report it.

Pinned by a case where the answer is known exactly: on a two-grain slab with a fully
transmitting boundary and mirror side walls every ballistic path needs two events
(boundary, then the x = L wall), so T = 1 without a limit and T = 0 with a limit of one
event -- were stopped paths counted as transmitted (as a lost particle is, RT:1833), T
would stay 1.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptrans import _lib, gen, geometry
from ptrans.constants import TRANS_VALUE
from test_extensions import _slab

OM = np.array([[3.0e13, 4.0e13]])


def _oracle_grid(geo, limit, np_=500, mfp=1.0e10, nangle=3, om=OM, seed=3):
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(om)
    O.set_event_limit(limit)
    try:
        return O.trace_grid(gdesc, mdesc, mfp, nangle, np_, seed, math=O.MATH_OWN)
    finally:
        O.set_event_limit(100_000_000)


def test_oracle_stopped_paths_are_not_transmitted():
    geo = _slab(TRANS_VALUE, 1.0)
    c, st = _oracle_grid(geo, 100_000_000)
    assert np.all(c[:, :, :2] == 500) and st[0] == 2 * 2 * 500
    c, st = _oracle_grid(geo, 1)
    assert not np.any(c) and st[0] == 2 * 2 * 500 and st[1] == 2 * 2 * 500
    gdesc, _g = geo.desc()
    one = np.ones((1, 1))
    mdesc, _m = geometry.modes_desc(np.array([[3e13]]), one, np.array([[1e10]]), one)
    prob, alias = O.build_alias(np.ones(1))
    O.set_event_limit(1)
    try:
        mp, mt, rec = O.trace_alias(gdesc, mdesc, np.array([[1e10]]), prob, alias, 5, 0, 300,
                                    records=True)
    finally:
        O.set_event_limit(100_000_000)
    assert int(mp.sum()) == 300 and int(mt.sum()) == 0
    assert np.all(rec["flags"] & 4) and np.all(rec["end_side"] == 0)


def _diffusive_film():
    return geometry.build_geometry(gen.film_holes(seed=2, nphonon=50, nangle=5, p_film=0.3))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 2])
def test_gpu_event_limit_matches_oracle(gpu_ctx, precision):
    """GPU counts equal the oracle's under the same limit: 0 on the slab with a limit of
    one event, and on a diffusive film (Lambda = 30 nm) with a limit of four events, where
    some paths end and others are stopped."""
    print("event limit, precision", precision, flush=True)
    om = np.linspace(2e12, 9e13, 6).reshape(2, 3)
    for geo, limit, mfp, np_ in ((_slab(TRANS_VALUE, 1.0), 1, 1.0e10, 500),
                                 (_diffusive_film(), 4, 30.0, 50)):
        oc, ost = _oracle_grid(geo, limit, np_=np_, mfp=mfp, nangle=5, om=om, seed=9)
        g = gpu_ctx.geometry(geo)
        m = gpu_ctx.modes(omega=om)
        gpu_ctx.set_seed(9)
        gpu_ctx.set_event_limit(limit)
        try:
            with pytest.warns(RuntimeWarning, match="event limit"):
                c, st = gpu_ctx.trace_grid(g, m, mfp, 5, np_, precision=precision)
        finally:
            gpu_ctx.set_event_limit(100_000_000)
        np.testing.assert_array_equal(c, oc)
        assert st["event_limited"] > 0 and st["paths"] == ost[0]
        g.close()
        m.close()


@pytest.mark.gpu
def test_gpu_event_limit_bound_and_lds_limit(gpu_ctx):
    """Limits above what a 32-bit draw counter allows are refused; an adjoint whose dynamic
    LDS fits but whose dynamic plus static LDS (16 KiB of lane counters) does not is
    refused with a message instead of failing at launch."""
    L = _lib.load()
    assert L.pt_ctx_set_event_limit(gpu_ctx.handle, 715827881) == 0
    assert L.pt_ctx_set_event_limit(gpu_ctx.handle, 715827882) == -1
    assert "draw counter" in _lib.last_error()
    gpu_ctx.set_event_limit(100_000_000)
    from test_batch import _modes
    omega, vel, mfp, weight, _, _ = _modes()
    geo = geometry.build_geometry(gen.film_holes(seed=2, nphonon=10, nangle=5))
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    nb = omega.shape[0]
    # dynamic LDS: nbatch * (3 + nb) doubles = 152,000 B (< 160 KiB), + 16 KiB static > 160 KiB
    nbatch = 152000 // (8 * (3 + nb))
    with pytest.raises(_lib.PtransError, match="LDS"):
        gpu_ctx.adjoint_alias(g, m, 1000, nbatch=nbatch)
