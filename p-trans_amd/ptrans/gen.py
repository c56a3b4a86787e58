"""Seeded generators of the benchmark and test configurations.

Every configuration is produced as an McpModel (the .mcp content P-TRANS reads,
src/RTMCphon.f90:252-1251) plus a synthetic bulk table, so the same input can be fed to
the reference (in the development container) and to this project's host code.

  nanowire        square wire along x, one Voronoi grain with six boundary facets
                  (the layout of Examples/01_square_nanowire.mcp)
  film_holes      film with z-aligned cylindrical holes (bricks + cylinder grains)
  inclusions      cube with sphere and z-cylinder inclusions (bricks)
  voronoi_cube    cube cut into Voronoi grains by half-space clipping
  tri_brick       a brick grain given as triangles (the `polyhedron` form, Examples/05)
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

from .mcp import FacetSpec, GrainSpec, McpModel, TransRow

Vec = Tuple[float, float, float]


def _box_faces(L: Sequence[float]):
    """The six faces of [0,Lx]x[0,Ly]x[0,Lz], counter-clockwise seen from outside,
    tagged -1..-6 (x-, x+, y-, y+, z-, z+)."""
    x, y, z = (float(v) for v in L)
    return [
        ([(0, 0, 0), (0, 0, z), (0, y, z), (0, y, 0)], -1),
        ([(x, 0, 0), (x, y, 0), (x, y, z), (x, 0, z)], -2),
        ([(0, 0, 0), (x, 0, 0), (x, 0, z), (0, 0, z)], -3),
        ([(x, y, 0), (0, y, 0), (0, y, z), (x, y, z)], -4),
        ([(0, 0, 0), (0, y, 0), (x, y, 0), (x, 0, 0)], -5),
        ([(0, y, z), (0, 0, z), (x, 0, z), (x, y, z)], -6),
    ]


def _clip_polyhedron(faces, n, d, tag, eps=1e-9):
    """Keep the part of a convex polyhedron with n.x <= d; the cut face gets `tag`."""
    out = []
    cap = []
    for verts, ftag in faces:
        res = []
        m = len(verts)
        for i in range(m):
            a, b = verts[i], verts[(i + 1) % m]
            da = n[0] * a[0] + n[1] * a[1] + n[2] * a[2] - d
            db = n[0] * b[0] + n[1] * b[1] + n[2] * b[2] - d
            if da <= eps:
                res.append(a)
                if abs(da) <= eps:
                    cap.append(a)
            if (da < -eps and db > eps) or (da > eps and db < -eps):
                t = da / (da - db)
                p = tuple(a[k] + t * (b[k] - a[k]) for k in range(3))
                res.append(p)
                cap.append(p)
        if len(res) >= 3:
            out.append((res, ftag))
    # cap polygon: unique points on the plane, ordered counter-clockwise about +n
    uniq: List[Vec] = []
    for p in cap:
        if all(max(abs(p[k] - q[k]) for k in range(3)) > 1e-7 for q in uniq):
            uniq.append(p)
    if len(uniq) >= 3:
        c = tuple(sum(p[k] for p in uniq) / len(uniq) for k in range(3))
        nn = np.asarray(n, dtype=float) / np.linalg.norm(n)
        u = np.asarray(uniq[0]) - np.asarray(c)
        u = u - nn * np.dot(u, nn)
        u /= np.linalg.norm(u)
        w = np.cross(nn, u)
        ang = [math.atan2(np.dot(np.asarray(p) - c, w), np.dot(np.asarray(p) - c, u)) for p in uniq]
        order = sorted(range(len(uniq)), key=lambda i: ang[i])
        out.append(([uniq[i] for i in order], tag))
    return out


def _clean_faces(faces, eps=1e-7):
    """Drop repeated consecutive vertices and degenerate faces."""
    res = []
    for verts, tag in faces:
        vv = []
        for p in verts:
            if not vv or max(abs(p[k] - vv[-1][k]) for k in range(3)) > eps:
                vv.append(p)
        if len(vv) > 1 and max(abs(vv[0][k] - vv[-1][k]) for k in range(3)) <= eps:
            vv.pop()
        if len(vv) >= 3:
            res.append((vv, tag))
    return res


def _one_material(kind="Value", value=0.0, p=0.0):
    return dict(materials=["Si"], nt=(1, 1), trans_rows=[TransRow("Si", "Si", kind, value, p)])


def nanowire(length=1000.0, width=100.0, height=100.0, p_wall=0.0, nphonon=100, nangle=11,
             mfp=None) -> McpModel:
    """C1-type wire: box length x width x height, one Voronoi grain whose six facets lie
    on the box faces (so they are not internal, RT:1216-1250); walls of specularity
    p_wall (CellSpecularity)."""
    faces = _box_faces((length, width, height))
    facets = [FacetSpec([tuple(map(float, v)) for v in vs], "Boundary", 0) for vs, _ in faces]
    g = GrainSpec((length / 2, width / 2, height / 2), "length", (25.0, 25.0, 25.0), "Si", facets)
    return McpModel(ncell=(1, 1, 1), lcell=(length, width, height),
                    cellspec=(p_wall,) * 6, nphonon=nphonon, nangle=nangle, mfp=mfp,
                    kind_grain="voronoi", grains=[g], **_one_material("Value", 0.0, 0.0))


def _place_disks(rng, n, rmin, rmax, lo, hi, others=(), log=False, max_tries=100000):
    """Non-overlapping circles/spheres: centres uniform in [lo+r, hi-r] per axis."""
    placed = []
    tries = 0
    while len(placed) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place inclusions without overlap")
        r = math.exp(rng.uniform(math.log(rmin), math.log(rmax))) if log else rng.uniform(rmin, rmax)
        c = [rng.uniform(lo[k] + r, hi[k] - r) for k in range(len(lo))]
        ok = True
        for (c2, r2, dims) in list(others) + [(p[0], p[1], len(lo)) for p in placed]:
            dd = sum((c[k] - c2[k]) ** 2 for k in range(min(dims, len(c))))
            if dd < (r + r2 + 1.0) ** 2:
                ok = False
                break
        if ok:
            placed.append((c, r))
    return placed


def film_holes(seed=2, lx=600.0, ly=600.0, h=200.0, radius=50.0, nholes=3, p_hole=0.3,
               p_film=0.0, p_side=1.0, nphonon=100, nangle=11, mfp=None) -> McpModel:
    """C2-type film: box lx x ly x h (transport along x, film faces z-/z+ of specularity
    p_film, in-plane y walls p_side), nholes z-cylinders of the given radius (vacuum:
    t = 0, specularity p_hole), centres seeded uniform without overlap and inside the
    cell.  The reference has no periodic boundary; p_side = 1 mirrors the cell."""
    rng = np.random.default_rng(seed)
    disks = _place_disks(rng, nholes, radius, radius, (0.0, 0.0), (lx, ly))
    grains = []
    for c, r in disks:
        grains.append(GrainSpec((c[0], c[1], h / 2), "cylinder", (2 * r, 2 * r, h), "Si",
                                [FacetSpec([], "Internal", 0)]))
    return McpModel(ncell=(1, 1, 1), lcell=(lx, ly, h),
                    cellspec=(1.0, 1.0, p_side, p_side, p_film, p_film),
                    nphonon=nphonon, nangle=nangle, mfp=mfp, kind_grain="bricks", grains=grains,
                    **_one_material("Value", 0.0, p_hole))


def inclusions(seed=3, L=1000.0, nsphere=8, ncyl=4, rs=(40.0, 120.0), rc=(30.0, 80.0),
               kind="Value", value=0.0, p=0.5, p_box=0.0, nphonon=100, nangle=11,
               mfp=None) -> McpModel:
    """C3-type cube: nsphere spheres (radius log-uniform in rs) and ncyl z-cylinders
    (radius uniform in rc), none overlapping and none touching the box faces.  One
    material pair law for every inclusion surface (the reference's specularity is per
    material pair, RT:1150)."""
    rng = np.random.default_rng(seed)
    cyl = _place_disks(rng, ncyl, rc[0], rc[1], (0.0, 0.0), (L, L))
    others = [(c, r, 2) for c, r in cyl]
    sph = _place_disks(rng, nsphere, rs[0], rs[1], (0.0, 0.0, 0.0), (L, L, L), others=others,
                       log=True)
    grains = []
    for c, r in sph:
        grains.append(GrainSpec(tuple(c), "sphere", (2 * r, 2 * r, 2 * r), "Si",
                                [FacetSpec([], "Internal", 0)]))
    for c, r in cyl:
        grains.append(GrainSpec((c[0], c[1], L / 2), "cylinder", (2 * r, 2 * r, L), "Si",
                                [FacetSpec([], "Internal", 0)]))
    return McpModel(ncell=(1, 1, 1), lcell=(L, L, L), cellspec=(p_box,) * 6,
                    nphonon=nphonon, nangle=nangle, mfp=mfp, kind_grain="bricks", grains=grains,
                    **_one_material(kind, value, p))


def voronoi_cube(seed=4, L=1000.0, ngrain=5, kind="Gamma", value=2.0, p=0.2, p_box=0.0,
                 nphonon=100, nangle=11, mfp=None) -> McpModel:
    """C4-type polycrystal: L^3 cube cut into ngrain Voronoi cells of seeded nucleation
    points by half-space clipping.  Facets on the box are 'Boundary' (not internal);
    shared facets name the neighbour grain.  Facets are counter-clockwise seen from
    outside their cell, so normals point out of the grain as the trace assumes."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0.15 * L, 0.85 * L, size=(ngrain, 3))
    grains = []
    for i in range(ngrain):
        faces = _box_faces((L, L, L))
        for j in range(ngrain):
            if j == i:
                continue
            n = tuple(float(v) for v in (pts[j] - pts[i]))
            d = float(np.dot(pts[j] - pts[i], (pts[i] + pts[j]) / 2))
            faces = _clean_faces(_clip_polyhedron(faces, n, d, j + 1))
        facets = []
        for verts, tag in faces:
            v = [tuple(float(c) for c in q) for q in verts]
            if tag < 0:
                facets.append(FacetSpec(v, "Boundary", 0))
            else:
                facets.append(FacetSpec(v, "Si", int(tag)))
        grains.append(GrainSpec(tuple(float(c) for c in pts[i]), "length", (25.0, 25.0, 25.0),
                                "Si", facets))
    return McpModel(ncell=(1, 1, 1), lcell=(L, L, L), cellspec=(p_box,) * 6,
                    nphonon=nphonon, nangle=nangle, mfp=mfp, kind_grain="voronoi", grains=grains,
                    **_one_material(kind, value, p))


def two_grain_slab(L=1000.0, W=200.0, a=400.0, kind="Value", value=1.0, p=1.0, p_side=1.0,
                   nphonon=100, nangle=3, mfp=None) -> McpModel:
    """Box L x W x W cut by the plane x = a into two Voronoi grains: grain 1 = [0, a] (the
    only incident grain), grain 2 = [a, L].  The shared facet carries the material pair's
    law and specularity; side walls y, z have specularity p_side (1: mirrors, which keep a
    path's |d_x|, so every path meets the boundary once from grain 1)."""
    box = (L, W, W)
    halves = [(_clip_polyhedron(_box_faces(box), (1.0, 0.0, 0.0), a, 2), (0.5 * a, 0.5 * W, 0.5 * W)),
              (_clip_polyhedron(_box_faces(box), (-1.0, 0.0, 0.0), -a, 1),
               (0.5 * (a + L), 0.5 * W, 0.5 * W))]
    grains = []
    for faces, gen in halves:
        facets = []
        for verts, tag in _clean_faces(faces):
            v = [tuple(float(c) for c in q) for q in verts]
            facets.append(FacetSpec(v, "Boundary", 0) if tag < 0 else FacetSpec(v, "Si", int(tag)))
        grains.append(GrainSpec(gen, "length", (25.0, 25.0, 25.0), "Si", facets))
    return McpModel(ncell=(1, 1, 1), lcell=box, cellspec=(1.0, 1.0, p_side, p_side, p_side, p_side),
                    nphonon=nphonon, nangle=nangle, mfp=mfp, kind_grain="voronoi", grains=grains,
                    **_one_material(kind, value, p))


def cube_triangles(lo: Sequence[float], hi: Sequence[float], nsub: int = 1, jitter: float = 0.0,
                   rng=None):
    """The surface of an axis-aligned brick as 12*nsub^2 triangles, counter-clockwise
    seen from outside; interior grid vertices optionally jittered within the face plane
    directions by up to `jitter` (edges and corners stay fixed so faces stay closed)."""
    tris = []
    lo = [float(v) for v in lo]
    hi = [float(v) for v in hi]
    for axis in range(3):
        for side in (0, 1):
            u_ax, v_ax = [(1, 2), (2, 0), (0, 1)][axis]
            grid = {}
            for a in range(nsub + 1):
                for b in range(nsub + 1):
                    p = [0.0, 0.0, 0.0]
                    p[axis] = hi[axis] if side else lo[axis]
                    p[u_ax] = lo[u_ax] + (hi[u_ax] - lo[u_ax]) * a / nsub
                    p[v_ax] = lo[v_ax] + (hi[v_ax] - lo[v_ax]) * b / nsub
                    if jitter > 0 and 0 < a < nsub and 0 < b < nsub and rng is not None:
                        p[u_ax] += rng.uniform(-jitter, jitter)
                        p[v_ax] += rng.uniform(-jitter, jitter)
                    grid[(a, b)] = tuple(p)
            for a in range(nsub):
                for b in range(nsub):
                    q00, q10, q11, q01 = grid[(a, b)], grid[(a + 1, b)], grid[(a + 1, b + 1)], grid[(a, b + 1)]
                    # (u, v, axis) is right-handed: counter-clockwise in (u, v) faces +axis
                    if side:
                        tris.append([q00, q10, q11]); tris.append([q00, q11, q01])
                    else:
                        tris.append([q00, q11, q10]); tris.append([q00, q01, q11])
    return tris


def tri_brick(L=(1000.0, 500.0, 500.0), lo=(300.0, 100.0, 100.0), hi=(700.0, 400.0, 400.0),
              nsub=2, seed=5, jitter=0.0, kind="Value", value=0.0, p=0.5, p_box=0.0,
              nphonon=100, nangle=11, mfp=None) -> McpModel:
    """A box of size L holding one brick grain given as triangles (the `polyhedron` form
    of Examples/05), material_to Internal (so the fill grain holds reversed copies)."""
    rng = np.random.default_rng(seed)
    tris = cube_triangles(lo, hi, nsub=nsub, jitter=jitter, rng=rng)
    facets = [FacetSpec(t, "Internal", 0) for t in tris]
    c = tuple((lo[k] + hi[k]) / 2 for k in range(3))
    g = GrainSpec(c, "polyhedron", tuple(hi[k] - lo[k] for k in range(3)), "Si", facets)
    return McpModel(ncell=(1, 1, 1), lcell=tuple(L), cellspec=(p_box,) * 6, nphonon=nphonon,
                    nangle=nangle, mfp=mfp, kind_grain="bricks", grains=[g],
                    **_one_material(kind, value, p))


def sc_lattice(seed=6, L=1000.0, k=4, m=2, jitter=5.0, kind="Gamma", value=2.0, p=0.3,
               p_box=0.0, nphonon=1, nangle=3, mfp=None) -> McpModel:
    """C5-type structure: the L^3 box split into a k x k x k simple-cubic lattice of
    grains whose faces are triangulated (each face m x m squares, two triangles each,
    12 k^3 m^2 triangles over all grains), on one shared vertex lattice jittered by up to
    `jitter` nm (box-face vertices only within their face, so the box stays closed; the
    triangles of neighbouring grains share vertices).  Triangles are counter-clockwise
    seen from outside their grain; faces on the box are 'Boundary', shared faces name the
    neighbour grain.  kind_grain 'mesh' (start grain from the incident facet containing
    the start point, no inside test; an extension)."""
    rng = np.random.default_rng(seed)
    n = k * m + 1
    h = L / (k * m)
    base = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1) * h
    jit = rng.uniform(-jitter, jitter, size=base.shape)
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1)
    jit[(idx == 0) | (idx == n - 1)] = 0.0
    V = base + jit

    def vert(a, b, c):
        return tuple(float(x) for x in V[a, b, c])

    def gid(i, j, l):
        return 1 + i + k * (j + k * l)

    grains = []
    for l in range(k):
        for j in range(k):
            for i in range(k):
                lo = (i * m, j * m, l * m)
                facets = []
                for axis in range(3):
                    u_ax, v_ax = [(1, 2), (2, 0), (0, 1)][axis]
                    for side in (0, 1):
                        cell = [i, j, l]
                        cell[axis] += 1 if side else -1
                        if 0 <= cell[axis] < k:
                            to, mat = gid(*cell), "Si"
                        else:
                            to, mat = 0, "Boundary"
                        for a in range(m):
                            for b in range(m):
                                def q(da, db):
                                    g = list(lo)
                                    g[axis] += m if side else 0
                                    g[u_ax] += a + da
                                    g[v_ax] += b + db
                                    return vert(*g)
                                q00, q10, q11, q01 = q(0, 0), q(1, 0), q(1, 1), q(0, 1)
                                tris = ([[q00, q10, q11], [q00, q11, q01]] if side else
                                        [[q00, q11, q10], [q00, q01, q11]])
                                facets.extend(FacetSpec(t, mat, to) for t in tris)
                c = tuple(float(x) for x in (np.array(lo) + m / 2.0) * h)
                grains.append(GrainSpec(c, "polyhedron", (k * h,) * 3, "Si", facets))
    return McpModel(ncell=(1, 1, 1), lcell=(L, L, L), cellspec=(p_box,) * 6, nphonon=nphonon,
                    nangle=nangle, mfp=mfp, kind_grain="mesh", grains=grains,
                    **_one_material(kind, value, p))


def sc_lattice_flat(seed=6, L=1000.0, k=4, m=2, jitter=5.0, trans_kind=2, trans_param=2.0,
                    p=0.3, p_box=0.0):
    """sc_lattice built directly as flattened arrays (geometry.FlatGeometry), vectorised,
    for structures of up to ~1e5 triangles: the same vertices, facet order, flags, laws
    and polygon set-up (bit-identical to building sc_lattice surface by surface)."""
    from . import geometry as G
    from .constants import GRAIN_MESH, SHAPE_POLYGON, TRANS_VALUE
    rng = np.random.default_rng(seed)
    n = k * m + 1
    h = L / (k * m)
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1)
    base = idx * h
    jit = rng.uniform(-jitter, jitter, size=base.shape)
    jit[(idx == 0) | (idx == n - 1)] = 0.0
    V = base + jit
    # facets of one grain, in sc_lattice's order: axis, side, a, b, two triangles
    tri_idx, sides, axes = [], [], []
    for axis in range(3):
        u_ax, v_ax = [(1, 2), (2, 0), (0, 1)][axis]
        for side in (0, 1):
            for a in range(m):
                for b in range(m):
                    def q(da, db):
                        g = [0, 0, 0]
                        g[axis] += m if side else 0
                        g[u_ax] += a + da
                        g[v_ax] += b + db
                        return g
                    q00, q10, q11, q01 = q(0, 0), q(1, 0), q(1, 1), q(0, 1)
                    tris = ([[q00, q10, q11], [q00, q11, q01]] if side else
                            [[q00, q11, q10], [q00, q01, q11]])
                    tri_idx.extend(tris)
                    sides.extend([side, side])
                    axes.extend([axis, axis])
    tri_idx = np.array(tri_idx)                  # (F, 3, 3) offsets within the grain
    sides, axes = np.array(sides), np.array(axes)
    F = len(tri_idx)
    grains = [(i, j, l) for l in range(k) for j in range(k) for i in range(k)]
    ng = len(grains)
    lo = np.array([[i * m, j * m, l * m] for (i, j, l) in grains])          # (ng, 3)
    vi = lo[:, None, None, :] + tri_idx[None]                               # (ng, F, 3, 3)
    verts = V[vi[..., 0], vi[..., 1], vi[..., 2]]                           # (ng, F, 3, 3)
    gcell = np.array(grains)                                                # (ng, 3)
    nb = np.repeat(gcell[:, None, :], F, axis=1).copy()
    step = np.where(sides == 1, 1, -1)
    nb[:, np.arange(F), axes] += step[None, :]
    inside = np.all((nb >= 0) & (nb < k), axis=2)
    to = 1 + nb[..., 0] + k * (nb[..., 1] + k * nb[..., 2])
    to = np.where(inside, to, 0)
    # box walls (the reference's setupBoxGrain) and the grains' triangles
    box = G.box_grain((L, L, L), (p_box,) * 6)
    nf = ng * F
    surf = np.zeros(6 + nf, dtype=G.SURFACE_DTYPE)
    vpool = [np.asarray(w.vertices, dtype=np.float64) for w in box.surfaces]
    off = 0
    for i, w in enumerate(box.surfaces):
        r = surf[i]
        r["kind"], r["nvertex"], r["vertex_offset"] = SHAPE_POLYGON, 4, off
        r["internal"], r["incident"] = 1, 0
        r["trans_kind"], r["trans_param"], r["specularity"] = TRANS_VALUE, 0.0, w.specularity
        r["center"], r["normal"] = w.center, w.normal
        r["tangent"] = [w.tangent[j][q] for j in range(3) for q in range(3)]
        off += 4
    tv = verts.reshape(nf, 3, 3)
    c, nrm, tang = G.setup_triangles(tv[:, 0], tv[:, 1], tv[:, 2])
    s = surf[6:]
    s["kind"], s["nvertex"] = SHAPE_POLYGON, 3
    s["vertex_offset"] = off + 3 * np.arange(nf)
    xs = tv[:, :, 0]
    incident = np.all(xs <= 0.0, axis=1)
    on_face = np.zeros(nf, dtype=bool)
    for ax in range(3):
        on_face |= np.all(tv[:, :, ax] <= 0.0, axis=1) | np.all(tv[:, :, ax] >= L, axis=1)
    internal = ~incident & ~on_face
    s["internal"], s["incident"] = internal, incident
    s["grain_inside"] = np.repeat(np.arange(1, ng + 1), F)
    s["grain_outside"] = to.reshape(nf)
    s["trans_kind"] = np.where(internal, trans_kind, TRANS_VALUE)
    s["trans_param"] = np.where(internal, trans_param, 0.0)
    s["specularity"] = np.where(internal, p, 0.0)
    s["center"], s["normal"], s["tangent"] = c, nrm, tang
    s["radius"] = 0.0
    vertices = np.concatenate([np.concatenate(vpool).reshape(-1), tv.reshape(-1)])
    goff = np.concatenate([[0, 6], 6 + F * np.arange(1, ng + 1)]).astype(np.int32)
    gcen = (lo + m / 2.0) * h
    ginc = np.concatenate([[0], (gcell[:, 0] == 0).astype(np.int32)]).astype(np.int32)
    gen_ = np.concatenate([[0.0, 0.0, 0.0], gcen.reshape(-1)]).astype(np.float64)
    arrays = dict(grain_offset=goff, grain_incident=ginc, generatrix=gen_, surfaces=surf,
                  vertices=vertices, nvertex=len(vertices) // 3)
    return G.FlatGeometry((L, L, L), (1.0, 1.0), GRAIN_MESH, ng, -1, 1, arrays)


def c5_structures(n=256, seed=5, triangles=(2.0e3, 1.0e5), k=4, L=1000.0, jitter=5.0,
                  trans_kind=2, trans_param=2.0, p=0.3):
    """C5 batch: n jittered k x k x k simple-cubic grain structures whose triangle counts
    (unique facets, 480 m^2 for k = 4) are log-uniform in `triangles`, each with its own
    jitter seed.  Returns (geometries, triangle counts)."""
    rng = np.random.default_rng(seed)
    faces = 3 * k * k * (k - 1) + 6 * k * k
    geos, tris = [], []
    for j in range(n):
        t = math.exp(rng.uniform(math.log(triangles[0]), math.log(triangles[1])))
        m = max(1, int(round(math.sqrt(t / (2.0 * faces)))))
        geos.append(sc_lattice_flat(seed=int(rng.integers(1 << 30)), L=L, k=k, m=m, jitter=jitter,
                                    trans_kind=trans_kind, trans_param=trans_param, p=p))
        tris.append(2 * faces * m * m)
    return geos, tris
