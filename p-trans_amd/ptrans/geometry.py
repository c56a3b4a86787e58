"""Geometry host side: builds the grain/surface model the trace reads, with the
semantics of loadGeometry (src/RTMCphon.f90:894-1251), setupBoxGrain (RT:1255-1332),
setupFillGrain (RT:1336-1398), setupSurfaceTransmissivity (RT:1402-1429) and
setupSurfaceGeometry (RT:1433-1491), and flattens it into the C-ABI's pt_geom_desc.

All arithmetic is Python float (IEEE double, no contraction), in the reference's
operation order, so normals and frames are bit-identical to the reference's.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import List, Optional

import numpy as np

from .constants import (GRAIN_BRICKS, GRAIN_MESH, GRAIN_VORONOI, SHAPE_CYLINDER, SHAPE_POLYGON,
                        SHAPE_SPHERE, TRANS_GAMMA, TRANS_TBC, TRANS_VALUE)
from .mcp import McpModel, trans_kind_id


def cross(a, b):
    """cross_product, RT:185-193."""
    return (a[1] * b[2] - a[2] * b[1],
            a[2] * b[0] - a[0] * b[2],
            a[0] * b[1] - a[1] * b[0])


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def setup_polygon(vertices):
    """setupSurfaceGeometry, polygon branch (RT:1440-1487): centroid, the averaged fan
    normal (not renormalised, RT:1459) and the tangent frame.  Returns
    (center, normal, tangent) with tangent as a list of 3 columns."""
    nv = len(vertices)
    if nv < 3:
        raise ValueError("a polygon needs at least 3 vertices")
    V = [tuple(float(c) for c in v) for v in vertices] + [tuple(float(c) for c in vertices[0])]
    c = [0.0, 0.0, 0.0]
    for iv in range(nv):
        c = [c[k] + V[iv][k] for k in range(3)]
    inv = 1.0 / nv
    center = tuple(inv * c[k] for k in range(3))
    n = [0.0, 0.0, 0.0]
    for iv in range(nv - 1):                           # iv = 1 .. nv-1 (1-based)
        a = [V[iv + 1][k] - V[iv][k] for k in range(3)]
        b = [V[iv + 2][k] - V[iv][k] for k in range(3)]
        cr = cross(a, b)
        ln = math.sqrt(dot(cr, cr))
        n = [n[k] + cr[k] / ln for k in range(3)]
    a = [V[nv][k] - V[nv - 1][k] for k in range(3)]
    b = [V[1][k] - V[nv - 1][k] for k in range(3)]
    cr = cross(a, b)
    ln = math.sqrt(dot(cr, cr))
    n = [n[k] + cr[k] / ln for k in range(3)]
    normal = tuple(inv * n[k] for k in range(3))
    col3 = normal
    if abs(normal[2]) == 1.0 or (normal[0] ** 2 == 0.0 and normal[1] ** 2 == 0.0):
        col2 = (0.0, 1.0, 0.0)
        col1 = (1.0, 0.0, 0.0)
    else:
        r = math.sqrt(normal[0] * normal[0] + normal[1] * normal[1])
        col2 = (-normal[1] / r, normal[0] / r, 0.0)
        col1 = cross(col2, col3)
    return center, normal, (col1, col2, col3)


@dataclasses.dataclass
class Surface:
    kind: int
    vertices: Optional[List[tuple]]
    internal: bool = True
    incident: bool = False
    grain_inside: int = 0
    grain_outside: int = 0
    imaterial1: int = 1
    imaterial2: int = 1
    trans_kind: int = TRANS_VALUE
    trans_param: float = 0.0
    specularity: float = 0.0
    center: tuple = (0.0, 0.0, 0.0)
    radius: float = 0.0
    normal: tuple = (0.0, 0.0, 0.0)
    tangent: tuple = ((0.0,) * 3,) * 3


@dataclasses.dataclass
class Grain:
    incident: bool = False
    generatrix: tuple = (0.0, 0.0, 0.0)
    surfaces: List[Surface] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class Geometry:
    box: tuple
    scope: tuple
    kind_grain: int
    ngrain: int
    igrain_fill: int
    nmaterial: int
    grains: List[Grain]
    cellspec: tuple = (1.0,) * 6

    # ---------------------------------------------------------------- flattening
    def flatten(self):
        """Arrays of pt_geom_desc (include/ptrans.h): grain CSR, incident flags,
        generatrices, pt_surface records and the vertex pool."""
        offs = [0]
        surfs = []
        verts = []
        for g in self.grains:
            for s in g.surfaces:
                voff = len(verts)
                if s.kind == SHAPE_POLYGON:
                    verts.extend(s.vertices)
                surfs.append((s, voff))
            offs.append(len(surfs))
        rec = np.zeros(len(surfs), dtype=SURFACE_DTYPE)
        for i, (s, voff) in enumerate(surfs):
            r = rec[i]
            r["kind"] = s.kind
            r["nvertex"] = len(s.vertices) if s.kind == SHAPE_POLYGON else 0
            r["vertex_offset"] = voff
            r["internal"] = int(s.internal)
            r["incident"] = int(s.incident)
            r["grain_inside"] = s.grain_inside
            r["grain_outside"] = s.grain_outside
            r["trans_kind"] = s.trans_kind
            r["trans_param"] = s.trans_param
            r["specularity"] = s.specularity
            r["center"] = s.center
            r["radius"] = s.radius
            r["normal"] = s.normal
            r["tangent"] = [s.tangent[j][i2] for j in range(3) for i2 in range(3)]
        return dict(
            grain_offset=np.asarray(offs, dtype=np.int32),
            grain_incident=np.asarray([int(g.incident) for g in self.grains], dtype=np.int32),
            generatrix=np.asarray([g.generatrix for g in self.grains], dtype=np.float64).reshape(-1),
            surfaces=rec,
            vertices=np.asarray(verts, dtype=np.float64).reshape(-1) if verts else np.zeros(3),
            nvertex=len(verts),
        )

    def desc(self):
        """A pt_geom_desc (ctypes) plus the arrays it points to (keep both alive)."""
        arr = self.flatten()
        d = PtGeomDesc()
        d.box = (ctypes.c_double * 3)(*self.box)
        d.scope = (ctypes.c_double * 2)(*self.scope)
        d.kind_grain = self.kind_grain
        d.ngrain = self.ngrain
        d.igrain_fill = self.igrain_fill
        d.nmaterial = self.nmaterial
        d.ngrain_total = len(self.grains)
        d.nsurface = len(arr["surfaces"])
        d.nvertex = arr["nvertex"]
        d.grain_offset = arr["grain_offset"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.grain_incident = arr["grain_incident"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.generatrix = arr["generatrix"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.surfaces = ctypes.c_void_p(arr["surfaces"].ctypes.data)
        d.vertices = arr["vertices"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return d, arr


SURFACE_DTYPE = np.dtype([
    ("kind", np.int32), ("nvertex", np.int32), ("vertex_offset", np.int64),
    ("internal", np.int32), ("incident", np.int32),
    ("grain_inside", np.int32), ("grain_outside", np.int32),
    ("trans_kind", np.int32), ("reserved", np.int32),
    ("trans_param", np.float64), ("specularity", np.float64),
    ("center", np.float64, 3), ("radius", np.float64),
    ("normal", np.float64, 3), ("tangent", np.float64, 9),
], align=True)
assert SURFACE_DTYPE.itemsize == 184


class PtGeomDesc(ctypes.Structure):
    _fields_ = [
        ("box", ctypes.c_double * 3), ("scope", ctypes.c_double * 2),
        ("kind_grain", ctypes.c_int32), ("ngrain", ctypes.c_int32),
        ("igrain_fill", ctypes.c_int32), ("nmaterial", ctypes.c_int32),
        ("ngrain_total", ctypes.c_int32), ("nsurface", ctypes.c_int32),
        ("nvertex", ctypes.c_int64),
        ("grain_offset", ctypes.POINTER(ctypes.c_int32)),
        ("grain_incident", ctypes.POINTER(ctypes.c_int32)),
        ("generatrix", ctypes.POINTER(ctypes.c_double)),
        ("surfaces", ctypes.c_void_p),
        ("vertices", ctypes.POINTER(ctypes.c_double)),
    ]


class PtModesDesc(ctypes.Structure):
    _fields_ = [
        ("nbranch", ctypes.c_int32), ("nkpoint", ctypes.c_int32),
        ("omega", ctypes.POINTER(ctypes.c_double)), ("vel", ctypes.POINTER(ctypes.c_double)),
        ("mfp", ctypes.POINTER(ctypes.c_double)), ("weight", ctypes.POINTER(ctypes.c_double)),
        ("max_omega", ctypes.c_double), ("temperature", ctypes.c_double),
    ]


def modes_desc(omega, vel=None, mfp=None, weight=None, max_omega=None, temperature=300.0):
    """A pt_modes_desc over (nb, nk) arrays (vel, mfp [nm], weight optional: the alias
    trace needs them); returns (desc, keepalive)."""
    omega = np.asarray(omega, dtype=np.float64)
    nb, nk = omega.shape
    flat = []
    ptrs = []
    for a in (omega, vel, mfp, weight):
        if a is None:
            ptrs.append(None)
            continue
        f = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F"))
        flat.append(f)
        ptrs.append(f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    d = PtModesDesc()
    d.nbranch, d.nkpoint = nb, nk
    d.omega, d.vel, d.mfp, d.weight = ptrs
    d.max_omega = float(np.max(omega)) if max_omega is None else float(max_omega)
    d.temperature = float(temperature)
    return d, flat


# ------------------------------------------------------------------------ builders

def _box_walls(bx, by, bz):
    """The six walls of setupBoxGrain in order x-, x+, y-, y+, z-, z+ with the vertex
    orders the reference ends up with (RT:1281-1320)."""
    return [
        [(0.0, 0.0, 0.0), (0.0, 0.0, bz), (0.0, by, bz), (0.0, by, 0.0)],
        [(bx, 0.0, 0.0), (bx, by, 0.0), (bx, by, bz), (bx, 0.0, bz)],
        [(0.0, 0.0, 0.0), (bx, 0.0, 0.0), (bx, 0.0, bz), (0.0, 0.0, bz)],
        [(bx, by, 0.0), (0.0, by, 0.0), (0.0, by, bz), (bx, by, bz)],
        [(0.0, 0.0, 0.0), (0.0, by, 0.0), (bx, by, 0.0), (bx, 0.0, 0.0)],
        [(0.0, by, bz), (0.0, 0.0, bz), (bx, 0.0, bz), (bx, by, bz)],
    ]


def box_grain(box, cellspec) -> Grain:
    """setupBoxGrain (RT:1255-1332): six internal polygon walls, grain_inside =
    grain_outside = 0, specularity cellspec(i), transmissivity VALUE 0."""
    g = Grain(incident=False, generatrix=(0.0, 0.0, 0.0))
    for i, w in enumerate(_box_walls(*box)):
        c, n, t = setup_polygon(w)
        g.surfaces.append(Surface(kind=SHAPE_POLYGON, vertices=w, internal=True, incident=False,
                                  grain_inside=0, grain_outside=0, trans_kind=TRANS_VALUE,
                                  trans_param=0.0, specularity=float(cellspec[i]),
                                  center=c, radius=0.0, normal=n, tangent=t))
    return g


def _material_tables(m: McpModel, tables=None):
    """param%kindTrans / transmissivity / specularity (RT:770-875), indexed
    [imaterial2][imaterial1] (1-based); row Nt2+1 is the boundary row (t = p = 0).  A TBC
    row (MW/m^2/K) becomes GammaDeterm's gamma (RT:858-870) from the bulk table of its
    material1, tables[material1]."""
    nt1, nt2 = m.nt
    kind = {}
    trans = {}
    spec = {}
    for i1 in range(1, nt1 + 1):
        kind[(nt2 + 1, i1)] = TRANS_VALUE
        trans[(nt2 + 1, i1)] = 0.0
        spec[(nt2 + 1, i1)] = 0.0
    for r in m.trans_rows:
        i1 = m.materials.index(r.material1) + 1
        i2 = m.materials.index(r.material2) + 1
        k = trans_kind_id(r.kind)
        kind[(i2, i1)] = k
        spec[(i2, i1)] = float(r.specularity)
        if k == TRANS_TBC:
            from .laws import gamma_determ
            if not tables or r.material1 not in tables:
                raise ValueError("the TBC row %s-%s needs the bulk table of %s (tables=...)"
                                 % (r.material1, r.material2, r.material1))
            trans[(i2, i1)] = gamma_determ(tables[r.material1], float(r.value) * 1.0e6,
                                           float(m.temperature))
            continue
        trans[(i2, i1)] = float(r.value)
    return kind, trans, spec


def _set_trans(s: Surface, kind: int, value: float):
    """setupSurfaceTransmissivity (RT:1402-1429), stored as a law, not a table."""
    s.trans_kind = kind
    s.trans_param = value


def build_geometry(m: McpModel, tables=None, setup_seed=None) -> Geometry:
    """loadGeometry (RT:894-1251) on a parsed .mcp.  tables: {material: anphon.Anphon},
    needed for TBC rows.  setup_seed: Philox seed of the FluctuationSpecularity jitter,
    needed when the .mcp sets it above 0."""
    if m.specfluc > 0.0 and setup_seed is None:
        raise ValueError("FluctuationSpecularity > 0 draws per-surface jitter: pass setup_seed")
    ndraw = 0
    box = m.box
    kind_grain = {"voronoi": GRAIN_VORONOI, "bricks": GRAIN_BRICKS, "mesh": GRAIN_MESH}[m.kind_grain]
    ngrain = len(m.grains)
    igrain_fill = ngrain + 1 if kind_grain == GRAIN_BRICKS else -1
    kind_t, trans_t, spec_t = _material_tables(m, tables)
    grains = [box_grain(box, m.cellspec)]
    for ig, gs in enumerate(m.grains, start=1):
        g = Grain(incident=False, generatrix=tuple(float(x) for x in gs.generatrix))
        if kind_grain != GRAIN_BRICKS or gs.shape in ("box", "length", "polyhedron"):
            kshape = SHAPE_POLYGON
            diameter = 0.0
        elif gs.shape == "sphere":
            kshape, diameter = SHAPE_SPHERE, float(gs.lengths[0])
        elif gs.shape == "cylinder":
            kshape, diameter = SHAPE_CYLINDER, float(gs.lengths[0])
        else:
            raise ValueError("unknown grain shape: %s" % gs.shape)
        im1 = m.materials.index(gs.material) + 1
        for f in gs.facets:
            s = Surface(kind=kshape, vertices=None, internal=True, incident=False,
                        grain_inside=ig, imaterial1=im1, center=g.generatrix,
                        radius=diameter * 0.5)
            if kshape == SHAPE_POLYGON:
                s.vertices = [tuple(float(c) for c in v) for v in f.vertices]
                s.center, s.normal, s.tangent = setup_polygon(s.vertices)
                xs = [v[0] for v in s.vertices]
                if all(x <= 0.0 for x in xs):                      # isIncidentSurface
                    g.incident = True
                    s.incident = True
                    s.internal = False
                if not _is_internal(s.vertices, box):              # isInternalSurface
                    s.internal = False
            s.grain_outside = int(f.grain_to)
            g.surfaces.append(s)
            if not s.internal:
                continue
            if f.material_to == "Internal":
                s.imaterial2 = 1
                s.grain_outside = igrain_fill
            else:
                s.imaterial2 = m.materials.index(f.material_to) + 1
            s.specularity = spec_t[(s.imaterial2, s.imaterial1)]
            if m.specfluc > 0.0:                                   # RT:1151-1157
                from .laws import fluctuate_specularity, setup_uniforms
                u = setup_uniforms(int(setup_seed), ndraw, 12)
                ndraw += 12
                s.specularity = fluctuate_specularity(s.specularity, float(m.specfluc), u)
            _set_trans(s, kind_t[(s.imaterial2, s.imaterial1)], trans_t[(s.imaterial2, s.imaterial1)])
        grains.append(g)
    if kind_grain == GRAIN_BRICKS:
        grains.append(_fill_grain(grains, ngrain, igrain_fill, kind_t, trans_t, spec_t))
    return Geometry(box=box, scope=(m.scopey, m.scopez), kind_grain=kind_grain, ngrain=ngrain,
                    igrain_fill=igrain_fill, nmaterial=len(m.materials), grains=grains,
                    cellspec=tuple(m.cellspec))


def _is_internal(verts, box) -> bool:
    """isInternalSurface (RT:1216-1250)."""
    for axis in range(3):
        c = [v[axis] for v in verts]
        if all(x <= 0.0 for x in c):
            return False
        if all(x >= box[axis] for x in c):
            return False
    return True


def _fill_grain(grains, ngrain, igrain_fill, kind_t, trans_t, spec_t) -> Grain:
    """setupFillGrain (RT:1336-1398): reversed copies of every internal brick surface."""
    fill = Grain(incident=False, generatrix=(0.0, 0.0, 0.0))
    for ig in range(1, ngrain + 1):
        for sj in grains[ig].surfaces:
            if not sj.internal:
                continue
            s = Surface(kind=sj.kind, vertices=None, internal=True, incident=False,
                        grain_inside=igrain_fill, imaterial1=1, imaterial2=sj.imaterial1,
                        center=sj.center, radius=sj.radius)
            if sj.kind == SHAPE_POLYGON:
                nv = len(sj.vertices)
                closed = list(sj.vertices) + [sj.vertices[0]]
                rev = [closed[nv + 1 - iv] for iv in range(1, nv + 2)]   # vertex(nv+2-iv)
                s.vertices = rev[:nv]
                s.center, s.normal, s.tangent = setup_polygon(s.vertices)
            s.grain_outside = sj.grain_inside
            s.specularity = spec_t[(s.imaterial2, s.imaterial1)]
            _set_trans(s, kind_t[(s.imaterial2, s.imaterial1)], trans_t[(s.imaterial2, s.imaterial1)])
            fill.surfaces.append(s)
    return fill


def inclusion_groups(geo: Geometry):
    """Specularity groups of a bricks geometry, one per brick: a brick's own surfaces
    (grain_inside = brick) and the fill grain's reversed copies of them (grain_outside =
    brick, setupFillGrain RT:1336-1398) form group brick-1; box walls and the rest get
    -1.  Returns (group per surface in descriptor order, number of groups)."""
    surf = geo.flatten()["surfaces"]
    gin = surf["grain_inside"].astype(np.int64)
    gout = surf["grain_outside"].astype(np.int64)
    nb = geo.ngrain
    own = (gin >= 1) & (gin <= nb)
    copy = ~own & (gout >= 1) & (gout <= nb) & (gin == geo.igrain_fill)
    group = np.full(len(surf), -1, dtype=np.int32)
    group[own] = gin[own] - 1
    group[copy] = gout[copy] - 1
    return group, nb


def group_specularity(specularity, group, values):
    """Per-surface specularities with every surface of group g set to values[g]."""
    p = np.array(specularity, dtype=np.float64, copy=True)
    group = np.asarray(group)
    values = np.asarray(values, dtype=np.float64)
    sel = group >= 0
    p[sel] = values[group[sel]]
    return p


def set_periodic(geo: Geometry, axes=(1,)) -> Geometry:
    """Make the box side walls of the given axes (1 = y, 2 = z) periodic (an extension:
    PT_TRANS_PERIODIC, the reference's box walls always reflect).  Returns geo."""
    from .constants import TRANS_PERIODIC
    walls = geo.grains[0].surfaces
    for a in axes:
        if a not in (1, 2):
            raise ValueError("only the y (1) and z (2) walls can be periodic")
        for w in (2 * a, 2 * a + 1):
            walls[w].trans_kind = TRANS_PERIODIC
            walls[w].trans_param = 0.0
    return geo


def set_lorentz_boundaries(geo: Geometry, rng, f0_thz=(3.0, 10.0)) -> dict:
    """Give every internal grain boundary (a facet shared by two grains; both copies) the
    Lorentzian law t = 1/(1 + (omega/omega0)^2) with omega0 = 2 pi f0, f0 uniform in f0_thz
    (an extension, PT_TRANS_LORENTZ).  Returns {(grain_a, grain_b): omega0}."""
    from .constants import TRANS_LORENTZ
    w0 = {}
    for gi, g in enumerate(geo.grains):
        if gi == 0:
            continue
        for s in g.surfaces:
            if not s.internal or s.grain_outside <= 0:
                continue
            key = (min(gi, s.grain_outside), max(gi, s.grain_outside))
            if key not in w0:
                w0[key] = rng.uniform(*f0_thz) * 1.0e12 * 2.0 * math.pi
            s.trans_kind = TRANS_LORENTZ
            s.trans_param = w0[key]
    return w0


class FlatGeometry:
    """A geometry held directly as the flattened arrays of pt_geom_desc (for meshes too
    large to build surface by surface); desc() and flatten() as Geometry's."""

    def __init__(self, box, scope, kind_grain, ngrain, igrain_fill, nmaterial, arrays):
        self.box, self.scope = tuple(box), tuple(scope)
        self.kind_grain, self.ngrain, self.igrain_fill = kind_grain, ngrain, igrain_fill
        self.nmaterial = nmaterial
        self._arr = arrays
        self.grains = [None] * (len(arrays["grain_offset"]) - 1)

    def flatten(self):
        return self._arr

    def desc(self):
        arr = self._arr
        d = PtGeomDesc()
        d.box = (ctypes.c_double * 3)(*self.box)
        d.scope = (ctypes.c_double * 2)(*self.scope)
        d.kind_grain = self.kind_grain
        d.ngrain = self.ngrain
        d.igrain_fill = self.igrain_fill
        d.nmaterial = self.nmaterial
        d.ngrain_total = len(arr["grain_offset"]) - 1
        d.nsurface = len(arr["surfaces"])
        d.nvertex = arr["nvertex"]
        d.grain_offset = arr["grain_offset"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.grain_incident = arr["grain_incident"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.generatrix = arr["generatrix"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.surfaces = ctypes.c_void_p(arr["surfaces"].ctypes.data)
        d.vertices = arr["vertices"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return d, arr


def setup_triangles(v0, v1, v2):
    """setup_polygon for many triangles at once (arrays (n, 3)), with the same operations
    in the same order, so the results are bit-identical: centroid, averaged fan normal and
    the tangent frame (n, 9), columns t1, t2, n."""
    def cr(a, b):
        return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                         a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                         a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)

    def unit(c):
        ln = np.sqrt(c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1] + c[:, 2] * c[:, 2])
        return c / ln[:, None]

    inv = 1.0 / 3
    z = np.zeros_like(v0)
    center = inv * (((z + v0) + v1) + v2)
    n = z + unit(cr(v1 - v0, v2 - v0))
    n = n + unit(cr(v2 - v1, v0 - v1))
    n = n + unit(cr(v0 - v2, v1 - v2))
    normal = inv * n
    tang = np.zeros((len(v0), 9))
    tang[:, 6:9] = normal
    axis = (np.abs(normal[:, 2]) == 1.0) | ((normal[:, 0] ** 2 == 0.0) & (normal[:, 1] ** 2 == 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.sqrt(normal[:, 0] * normal[:, 0] + normal[:, 1] * normal[:, 1])
        c2 = np.stack([-normal[:, 1] / r, normal[:, 0] / r, np.zeros(len(v0))], axis=1)
    c1 = cr(c2, normal)
    c2[axis] = (0.0, 1.0, 0.0)
    c1[axis] = (1.0, 0.0, 0.0)
    tang[:, 0:3] = c1
    tang[:, 3:6] = c2
    return center, normal, tang
