"""Start-grain map of Voronoi geometries.  The device skips both start-grain searches
(findNearestIncidentGrain, RT:2100-2118, and the isInsideGrain check of the particle loop,
RT:1673-1682) for a start point in a cell of the start face that the host certified: on the
cell's corners one incident generatrix is nearest by a margin, every facet test of that grain
passes by a margin and every other grain has a facet that fails by a margin.  Elsewhere the
searches run as before.  Random start points over several Voronoi cubes (cells of either
kind, and the cells' edges) must give the oracle's start grain, exit, event count and draw
count path for path, in FP64 and in the FP32-filter precision.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptrans import gen, geometry

OM = np.linspace(2e12, 9e13, 6).reshape(2, 3)
CASES = [(4, 5), (11, 20), (23, 60)]


def _geo(seed, ngrain):
    m = gen.voronoi_cube(seed=seed, L=1000.0, ngrain=ngrain, kind="Gamma", value=2.0, p=0.3,
                         nphonon=2000, nangle=5)
    return geometry.build_geometry(m)


@pytest.mark.gpu
@pytest.mark.parametrize("seed, ngrain", CASES)
@pytest.mark.parametrize("precision", [0, 2])
def test_gpu_start_grains_equal_oracle(gpu_ctx, seed, ngrain, precision):
    geo = _geo(seed, ngrain)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(OM)
    n = 20000
    orec = O.replay(gdesc, mdesc, 1.0e10, 5, 2000, 31, 0, n, math=O.MATH_OWN)
    gpu_ctx.set_seed(31)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=OM)
    rec = gpu_ctx.replay(g, m, 1.0e10, 5, 2000, 0, n, precision=precision)
    grains = np.unique(orec["start_grain"])
    print("start grains", grains, "mismatches", int(np.sum(orec["flags"] & 2)), flush=True)
    assert len(grains) >= 3
    for f in ("start_grain", "flags", "end_side", "events", "draws"):
        np.testing.assert_array_equal(rec[f], orec[f], err_msg=f)
    np.testing.assert_array_equal(rec["pos"], orec["pos"])
