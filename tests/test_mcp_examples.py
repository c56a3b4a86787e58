"""The host's .mcp reader and geometry build (mcp.read_mcp + geometry.build_geometry)
against the reference's own loadMCP (RT:252-1251) on the reference's own inputs,
Examples/01-05: every parameter the trace reads and every surface record, bit for bit.

The fixtures tests/golden/examples/*.json hold what the unmodified reference built
(oracle/replay's `dump` mode, tests/golden/make_golden.py examples).  The .mcp inputs are
the reference's files, read where they lie: this module runs in the development
container, where /root/reference exists, and is skipped elsewhere.

What this pins: first-match keys (Examples/01 sets `temp` twice, RT:212-231), defaults,
positional grain blocks (RT:970-1164), radius = half the first number of a sphere or
cylinder line (RT:999-1015), the `polyhedron ... .STL` line of Examples/05 (whose facets
are inline, RT:1016-1019), incident/internal classification (RT:1104-1112, 1206-1250),
material_to / grain_to (RT:1118-1146), specularity and laws per material pair
(RT:1150, 1160-1162), the box grain (RT:1255-1332) and the fill grain's reversed copies
(RT:1336-1398), and setupSurfaceGeometry's centroid, normal and frame (RT:1433-1491).
"""
import json
import os

import numpy as np
import pytest

from _golden import GOLDEN
from ptrans import geometry, mcp
from ptrans.constants import SHAPE_POLYGON

EXAMPLES = "/root/reference/Examples"
NAMES = sorted(f[:-5] for f in os.listdir(os.path.join(GOLDEN, "examples")) if f.endswith(".json"))


def _load(name):
    with open(os.path.join(GOLDEN, "examples", name + ".json")) as f:
        return json.load(f)


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="the reference's Examples/ are read "
                    "where they lie (development container only)")
@pytest.mark.parametrize("name", NAMES)
def test_examples_parse_like_the_reference(name):
    ref = _load(name)
    model = mcp.read_mcp(os.path.join(EXAMPLES, name + ".mcp"))
    geo = geometry.build_geometry(model)
    p = ref["param"]
    assert list(geo.box) == p["box"]
    assert [model.scopey, model.scopez] == p["scope"]
    assert model.temperature == p["temperature"]
    assert model.nphonon == p["nphonon"] and model.nangle == p["nangle"]
    assert model.mean_free_paths() == ref["mfp"]
    assert list(model.cellspec) == ref["cellspec"]
    assert model.specfluc == p["specfluc"]
    assert geo.kind_grain == p["kind_grain"] and geo.nmaterial == p["nmaterial"]
    lo, hi, ibox, ifill = ref["grain_range"]
    assert (lo, ibox) == (0, 0) and hi == len(geo.grains) - 1
    assert geo.igrain_fill == (ifill if geo.kind_grain == 2 else -1)
    for g, rg in zip(geo.grains, ref["grains"]):
        ig = rg["index"]
        assert len(g.surfaces) == rg["nsurface"], ig
        # setupFillGrain sets neither the fill grain's incident flag nor its generatrix
        # (RT:1336-1398; both undefined, and the start-grain searches loop over 1..ngrain)
        if 0 < ig and ig != ifill:
            assert int(g.incident) == rg["incident"], ig
            assert list(g.generatrix) == rg["generatrix"], ig
        for k, (s, rs) in enumerate(zip(g.surfaces, rg["surfaces"])):
            where = "grain %d surface %d" % (ig, k + 1)
            assert s.kind == rs["kind"], where
            assert int(s.internal) == rs["internal"] and int(s.incident) == rs["incident"], where
            assert s.grain_inside == rs["grain_inside"], where
            if s.internal:
                # set only for surfaces the trace can hit (RT:1118-1146)
                assert s.grain_outside == rs["grain_outside"], where
                if ig > 0:          # setupBoxGrain sets no materials (RT:1255-1332)
                    assert (s.imaterial1, s.imaterial2) == (rs["imaterial1"], rs["imaterial2"]), where
                assert s.trans_kind == rs["trans_kind"], where
                assert s.specularity == rs["specularity"], where
                assert s.trans_param == rs["law"], where
            if s.kind == SHAPE_POLYGON:
                assert [list(v) for v in s.vertices] == rs["vertices"], where
                assert list(s.center) == rs["center"], where
                assert list(s.normal) == rs["normal"], where
                assert [x for col in s.tangent for x in col] == rs["tangent"], where
            else:
                assert s.radius == rs["radius"] and list(s.center) == rs["center"], where


def test_example_fixtures_cover_every_shape_and_grain_kind():
    """The five fixtures exercise polygons, spheres and cylinders, Voronoi and bricks grains,
    incident and internal facets and the fill grain."""
    kinds, grain_kinds, fill = set(), set(), False
    for name in NAMES:
        ref = _load(name)
        grain_kinds.add(ref["param"]["kind_grain"])
        fill |= ref["grain_range"][3] > 0 and ref["param"]["kind_grain"] == 2
        for g in ref["grains"]:
            for s in g["surfaces"]:
                kinds.add(s["kind"])
    assert kinds == {1, 2, 3} and grain_kinds == {1, 2} and fill and len(NAMES) == 5
