"""Start-grain map.  The device skips both start-grain searches (findNearestIncidentGrain,
RT:2100-2118, or findIncidentCollisionGrain, RT:2131-2163, and the isInsideGrain check of the
particle loop, RT:1673-1682) for a start point in a cell of the start face that the host
certified: Voronoi, on the cell's corners one incident generatrix is nearest by a margin;
bricks with no incident grain, the fill grain; then every surface test of that grain passes
by a margin on the whole cell and every other grain has a surface that fails by a margin
(polygons by their affine plane test, spheres and z-cylinders by distance bounds over the
cell).  Elsewhere the searches run as before.  Random start points over Voronoi cubes and
over brick geometries with spherical and cylindrical grains must give the oracle's start
grain, exit, event count and draw count path for path, in FP64 and in the FP32-filter
precision.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from ptrans import gen, geometry

OM = np.linspace(2e12, 9e13, 6).reshape(2, 3)
CASES = ["voronoi5", "voronoi20", "voronoi60", "film_holes", "inclusions", "inclusions_at_x0"]


def _geo(case):
    if case.startswith("voronoi"):
        ngrain = int(case[7:])
        m = gen.voronoi_cube(seed={5: 4, 20: 11, 60: 23}[ngrain], L=1000.0, ngrain=ngrain,
                             kind="Gamma", value=2.0, p=0.3, nphonon=2000, nangle=5)
    elif case == "film_holes":
        m = gen.film_holes(seed=2, nphonon=2000, nangle=5)
    elif case == "inclusions":
        m = gen.inclusions(seed=3, nphonon=2000, nangle=5)
    else:
        # spheres cut by the start face: start points inside them are certain mismatches
        # (the particle loop's EXIT) and cells across their rims are left to the searches
        m = gen.inclusions(seed=3, nphonon=2000, nangle=5)
        m.grains = [dataclasses.replace(g, generatrix=(0.25 * g.lengths[0],) + tuple(g.generatrix[1:]))
                    if g.shape == "sphere" else g for g in m.grains]
    return geometry.build_geometry(m)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("precision", [0, 2])
def test_gpu_start_grains_equal_oracle(gpu_ctx, case, precision):
    geo = _geo(case)
    gdesc, _g = geo.desc()
    mdesc, _m = geometry.modes_desc(OM)
    n = 20000
    orec = O.replay(gdesc, mdesc, 1.0e10, 5, 2000, 31, 0, n, math=O.MATH_OWN)
    gpu_ctx.set_seed(31)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=OM)
    rec = gpu_ctx.replay(g, m, 1.0e10, 5, 2000, 0, n, precision=precision)
    grains = np.unique(orec["start_grain"])
    print(case, "start grains", grains, "mismatches", int(np.sum(orec["flags"] & 2)), flush=True)
    if case.startswith("voronoi"):
        assert len(grains) >= 3
    if case == "inclusions_at_x0":
        assert np.sum(orec["flags"] & 2) > 0
    for f in ("start_grain", "flags", "end_side", "events", "draws"):
        np.testing.assert_array_equal(rec[f], orec[f], err_msg=f)
    np.testing.assert_array_equal(rec["pos"], orec["pos"])
