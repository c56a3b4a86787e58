"""Geometry kernels against the reference's own module routines (golden file written by
oracle/replay/driver.f90 'geom' mode): setupSurfaceGeometry (RT:1433-1491),
calcSurfaceDistancePolygon/Sphere/Cylinder (RT:2907-3088), isInsidePolygon
(RT:2961-2983), calcSurfaceNormal* (RT:3119-3179), calcSurfaceRotation* (RT:3181-3301).
Bit-exact, for the oracle and for the Python host's setup_polygon.  Plus closed forms."""
import ctypes
import math
import os

import numpy as np
import pytest

from _golden import GOLDEN
from oracle import oracle as O
from ptrans import geometry


def _tokens(path):
    with open(path) as f:
        return f.read().split()


def _cases():
    t = _tokens(os.path.join(GOLDEN, "geom_cases.txt"))
    i, out = 0, []
    while i < len(t):
        tag = t[i]; i += 1
        if tag == "POLY":
            nv = int(t[i]); i += 1
            v = np.array([float(x) for x in t[i:i + 3 * nv]]).reshape(nv, 3); i += 3 * nv
            rest = np.array([float(x) for x in t[i:i + 9]]); i += 9
            out.append((tag, dict(v=v, rs=rest[0:3], nd=rest[3:6], p=rest[6:9])))
        else:
            a = np.array([float(x) for x in t[i:i + 11]]); i += 11
            out.append((tag, dict(rs=a[0:3], nd=a[3:6], c=a[6:9], r=a[9], boxz=a[10])))
    return out


def _refs():
    t = _tokens(os.path.join(GOLDEN, "geom_ref.txt"))
    i, out = 0, []
    f = lambda k: [float(x) for x in t[i:i + k]]
    while i < len(t):
        tag = t[i]; i += 1
        if tag == "POLY":
            rec = dict(center=f(3)); i += 3
            rec["normal"] = f(3); i += 3
            rec["tangent"] = f(9); i += 9
            rec["hit"] = int(t[i]); rec["l"] = float(t[i + 1]); i += 2
            rec["inside"] = int(t[i]); i += 1
        else:
            rec = dict(hit=int(t[i]), l=float(t[i + 1])); i += 2
            rec["p"] = f(3); i += 3
            rec["n"] = f(3); i += 3
            rec["rot"] = f(9); i += 9
        out.append((tag, rec))
    return out


CASES = list(zip(_cases(), _refs()))


def _same(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.array_equal(a, b, equal_nan=True)


_KEEP = []


def _p(a):
    """Pointer to a float64 copy of a; the copy is kept alive for the test session."""
    arr = np.ascontiguousarray(a, dtype=np.float64)
    _KEEP.append(arr)
    return arr.ctypes.data


def test_golden_file_is_complete():
    assert len(CASES) == 400
    assert [c[0][0] for c in CASES] == [c[1][0] for c in CASES]


@pytest.mark.parametrize("k", range(0, 400, 3))
def test_polygon_setup_distance_inside(k):
    (tag, inp), (rtag, ref) = CASES[k]
    assert tag == rtag == "POLY"
    L = O.lib()
    v = np.ascontiguousarray(inp["v"])
    nv = len(v)
    c, n, t = np.zeros(3), np.zeros(3), np.zeros(9)
    L.oracle_setup_polygon(nv, v.ctypes.data, c.ctypes.data, n.ctypes.data, t.ctypes.data)
    assert _same(c, ref["center"]) and _same(n, ref["normal"]) and _same(t, ref["tangent"])
    pc, pn, pt = geometry.setup_polygon([tuple(x) for x in v])
    assert _same(pc, ref["center"]) and _same(pn, ref["normal"])
    assert _same([x for col in pt for x in col], ref["tangent"])
    l = np.zeros(1)
    hit = L.oracle_dist_polygon(_p(inp["rs"]), _p(inp["nd"]), v.ctypes.data, nv, n.ctypes.data,
                                l.ctypes.data)
    assert hit == ref["hit"]
    assert l[0] == ref["l"]
    assert L.oracle_inside_polygon(v.ctypes.data, nv, _p(inp["p"])) == ref["inside"]


@pytest.mark.parametrize("k", [k for k in range(400) if k % 3 != 0])
def test_sphere_cylinder_distance_normal_frame(k):
    (tag, inp), (rtag, ref) = CASES[k]
    assert tag == rtag
    L = O.lib()
    l = np.zeros(1)
    if tag == "SPH":
        hit = L.oracle_dist_sphere(_p(inp["rs"]), _p(inp["nd"]), _p(inp["c"]), float(inp["r"]), l.ctypes.data)
        kind = 2
    else:
        hit = L.oracle_dist_cylinder(_p(inp["rs"]), _p(inp["nd"]), _p(inp["c"]), float(inp["r"]),
                                     float(inp["boxz"]), l.ctypes.data)
        kind = 3
    assert hit == ref["hit"]
    if hit:
        assert l[0] == ref["l"]
        p = inp["rs"] + inp["nd"] * l[0]
    else:
        p = inp["rs"].copy()
    assert _same(p, ref["p"])
    n, rot = np.zeros(3), np.zeros(9)
    L.oracle_surface_normal(kind, _p(p), _p(inp["c"]), float(inp["r"]), None, n.ctypes.data)
    L.oracle_surface_rotation(kind, _p(p), _p(inp["c"]), None, rot.ctypes.data)
    assert _same(n, ref["n"])
    assert _same(rot, ref["rot"])


def test_closed_forms():
    """Unit sphere at the origin from (-3,0,0) along +x: hit at distance 2, outward
    normal flipped inward (|h-c|^2 > r^2 is false on the surface, so +h); a z-cylinder of
    radius 1: same in the xy plane; a ray along the axis misses (0/0); a square in the
    z = 0 plane is hit only from below (one-sided)."""
    L = O.lib()
    l = np.zeros(1)
    assert L.oracle_dist_sphere(_p([-3, 0, 0]), _p([1, 0, 0]), _p([0, 0, 0]), 1.0, l.ctypes.data) == 1
    assert l[0] == 2.0
    assert L.oracle_dist_sphere(_p([-3, 0, 0]), _p([-1, 0, 0]), _p([0, 0, 0]), 1.0, l.ctypes.data) == 0
    assert L.oracle_dist_sphere(_p([0.5, 0, 0]), _p([1, 0, 0]), _p([0, 0, 0]), 1.0, l.ctypes.data) == 1
    assert l[0] == 0.5
    assert L.oracle_dist_cylinder(_p([-3, 0, 5]), _p([1, 0, 0]), _p([0, 0, 0]), 1.0, 10.0, l.ctypes.data) == 1
    assert l[0] == 2.0
    assert L.oracle_dist_cylinder(_p([-3, 0, 5]), _p([1, 0, 0]), _p([0, 0, 0]), 1.0, 4.0, l.ctypes.data) == 0
    assert L.oracle_dist_cylinder(_p([0.2, 0, 5]), _p([0, 0, 1]), _p([0, 0, 0]), 1.0, 10.0, l.ctypes.data) == 0
    sq = np.ascontiguousarray([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=float)
    c, n, t = np.zeros(3), np.zeros(3), np.zeros(9)
    L.oracle_setup_polygon(4, sq.ctypes.data, c.ctypes.data, n.ctypes.data, t.ctypes.data)
    assert list(n) == [0.0, 0.0, 1.0] and list(c) == [0.5, 0.5, 0.0]
    assert L.oracle_dist_polygon(_p([0.5, 0.5, -2]), _p([0, 0, 1]), sq.ctypes.data, 4, n.ctypes.data, l.ctypes.data) == 1
    assert l[0] == 2.0
    assert L.oracle_dist_polygon(_p([0.5, 0.5, 2]), _p([0, 0, -1]), sq.ctypes.data, 4, n.ctypes.data, l.ctypes.data) == 0
    assert L.oracle_dist_polygon(_p([1.5, 0.5, -2]), _p([0, 0, 1]), sq.ctypes.data, 4, n.ctypes.data, l.ctypes.data) == 0
    # on-edge counts as inside (RT:2950)
    assert L.oracle_inside_polygon(sq.ctypes.data, 4, _p([1.0, 0.5, 0.0])) == 1
