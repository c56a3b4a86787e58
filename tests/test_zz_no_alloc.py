"""pt_trace_alias_noalloc, the allocation-free alias trace (for callers that cannot
allocate device memory): through the C-ABI it returns exactly what pt_trace_alias returns
for the same seed -- per-mode path and transmitted counts, path and event statistics --
on two calls in a row, the second with a smaller geometry, table and path count (the
shared scratch keeps the first call's contents, so the second checks that what must start
at zero does).  The caller's tally arrays are added to, as by pt_trace_alias."""
import ctypes

import numpy as np
import pytest

from ptrans import _lib, gen, geometry
from ptrans.anphon import synthetic_table

M_TWOPI = 3.14159265358979323846 * 2.0


def _table(nb, nk, seed):
    t = synthetic_table(nb, nk, seed)
    return t.omega_cm() * 0.03 * M_TWOPI * 1.0e12, t.vel, t.mfp_nm, t.weight


def _call(ctx, g, m, npaths, precision, scale, prefill):
    nm = m.nbranch * m.nkpoint
    mp = np.array(prefill[0], dtype=np.int64)
    mt = np.array(prefill[1], dtype=np.int64)
    st = _lib.PtStats()
    sc = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
    rc = _lib.load().pt_trace_alias_noalloc(ctx.handle, g.handle, m.handle, int(npaths),
                                            int(precision), sc.ctypes.data if sc is not None else None,
                                            mp.ctypes.data, mt.ctypes.data, ctypes.byref(st))
    _lib.check(rc)
    assert mp.size == nm
    return mp, mt, st.as_dict()


@pytest.mark.gpu
def test_gpu_noalloc_equals_trace_alias_on_two_calls(gpu_ctx):
    cases = [
        # inclusions cube, 6 x 2000 modes, branch scales, FP32 trace
        (geometry.build_geometry(gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.5,
                                                nphonon=1, nangle=3)),
         _table(6, 2000, 20), 300_000, 1, np.linspace(0.8, 1.3, 6)),
        # then a smaller one: Voronoi polycrystal, 3 x 500 modes, FP32 filter (FP64 bits)
        (geometry.build_geometry(gen.voronoi_cube(seed=4, kind="Gamma", value=1.5, p=0.3,
                                                  nphonon=1, nangle=3)),
         _table(3, 500, 21), 40_000, 2, None),
    ]
    for i, (geo, tab, npaths, prec, scale) in enumerate(cases):
        print("no-alloc call %d: %d paths, %d modes, precision %d" % (i, npaths, tab[0].size, prec),
              flush=True)
        omega, vel, mfp, weight = tab
        g = gpu_ctx.geometry(geo)
        m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
        gpu_ctx.set_seed(1000 + i)
        ref_p, ref_t, ref_st = gpu_ctx.trace_alias(g, m, npaths, branch_scale=scale, precision=prec)
        nm = omega.size
        pre = (np.arange(nm, dtype=np.int64), 3 * np.arange(nm, dtype=np.int64))
        mp, mt, st = _call(gpu_ctx, g, m, npaths, prec, scale, pre)
        np.testing.assert_array_equal(mp - pre[0], ref_p.reshape(-1, order="F"))
        np.testing.assert_array_equal(mt - pre[1], ref_t.reshape(-1, order="F"))
        # results of the paths (retests and fallbacks count work, which depends on which
        # paths a wave serves alone at the end of the launch, so they vary run to run)
        for k in ("paths", "events", "draws", "lost", "start_mismatch", "event_limited"):
            assert st[k] == ref_st[k], k
        assert st["paths"] == npaths
        # the Python wrapper's switch reaches the same entry point
        wp, wt, _ = gpu_ctx.trace_alias(g, m, npaths, branch_scale=scale, precision=prec, noalloc=True)
        np.testing.assert_array_equal(wp, ref_p)
        np.testing.assert_array_equal(wt, ref_t)
        g.close()
        m.close()


@pytest.mark.gpu
def test_gpu_noalloc_refuses_tables_beyond_its_scratch(gpu_ctx):
    geo = geometry.build_geometry(gen.nanowire(p_wall=0.5, nphonon=1, nangle=3))
    g = gpu_ctx.geometry(geo)
    nb, nk = 2, 40000                                   # 80,000 modes > 65,536
    omega = np.linspace(1e12, 9e13, nb * nk).reshape(nb, nk)
    one = np.ones((nb, nk))
    m = gpu_ctx.modes(omega=omega, vel=one * 5000.0, mfp=one * 100.0, weight=one)
    with pytest.raises(_lib.PtransError, match="scratch"):
        _call(gpu_ctx, g, m, 10, 0, None, (np.zeros(nb * nk), np.zeros(nb * nk)))
    g.close()
    m.close()
