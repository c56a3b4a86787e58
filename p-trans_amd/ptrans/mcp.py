"""The .mcp input file: a writer for the seeded configurations and a reader with the
semantics of loadMCP (src/RTMCphon.f90:252-1251).

Reader rules kept from the reference:
  - findKey (RT:212-231): a key matches the first whitespace token of a non-empty line;
    the FIRST matching line wins; the value is on the following line;
  - defaults of optional keys (CellSpecularity 1, FluctuationSpecularity 0, Nangle 91,
    ScopeY/ScopeZ 1, MeanFreePath absent -> Matthiessen with 1e10, Trace 1000);
  - box = L*cell * N*cell (RT:338-380);
  - the grain blocks after the Ng value line are positional (RT:970-1164); a sphere's or
    cylinder's diameter is the first number of the line after its shape word (RT:999-1014).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Tuple

from .constants import TRANS_GAMMA, TRANS_TBC, TRANS_VALUE


@dataclasses.dataclass
class FacetSpec:
    vertices: List[Tuple[float, float, float]]
    material_to: str
    grain_to: int


@dataclasses.dataclass
class GrainSpec:
    generatrix: Tuple[float, float, float]
    shape: str                    # box | length | sphere | cylinder | polyhedron
    lengths: Tuple[float, ...]    # the line after the shape word (diameter first)
    material: str
    facets: List[FacetSpec]       # polygon grains; one entry (no vertices) otherwise


@dataclasses.dataclass
class TransRow:
    material1: str
    material2: str
    kind: str                     # Value | Gamma | TBC
    value: float                  # t, gamma or TBC [MW/m^2/K]
    specularity: float


@dataclasses.dataclass
class McpModel:
    ncell: Tuple[int, int, int] = (1, 1, 1)
    lcell: Tuple[float, float, float] = (100.0, 100.0, 100.0)
    cellspec: Tuple[float, ...] = (1.0,) * 6
    specfluc: float = 0.0
    temperature: float = 300.0
    nphonon: int = 100
    nangle: int = 91
    scopey: float = 1.0
    scopez: float = 1.0
    mfp: Optional[tuple] = None   # None, (value,), or (min, max, n, 'linear'|'log')
    trace: int = 1000
    materials: List[str] = dataclasses.field(default_factory=lambda: ["Si"])
    nt: Tuple[int, int] = (1, 1)
    trans_rows: List[TransRow] = dataclasses.field(default_factory=list)
    kind_grain: str = "bricks"
    grains: List[GrainSpec] = dataclasses.field(default_factory=list)

    @property
    def box(self) -> Tuple[float, float, float]:
        return tuple(float(self.lcell[i]) * self.ncell[i] for i in range(3))

    def mean_free_paths(self) -> List[float]:
        """param%vMeanFreePath (RT:507-546)."""
        if self.mfp is None:
            return [1.0e10]
        if len(self.mfp) == 1:
            return [float(self.mfp[0])]
        lo, hi, n, scale = self.mfp
        if scale == "linear":
            return [lo + (hi - lo) / (n - 1) * (s - 1) for s in range(1, n + 1)]
        return [math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) / (n - 1) * (s - 1))
                for s in range(1, n + 1)]


def _fmt(x: float) -> str:
    return "%.17g" % x


def write_mcp(path: str, m: McpModel) -> None:
    """Write a file that loadMCP (RT:252) and read_mcp() both read back to `m`."""
    L = []
    L += ["# ptrans seeded configuration"]
    for key, v in zip(("NXcell", "NYcell", "NZcell"), m.ncell):
        L += [key, "\t%d" % v]
    for key, v in zip(("LXcell", "LYcell", "LZcell"), m.lcell):
        L += [key, "\t" + _fmt(v)]
    L += ["", "## Raytracing"]
    L += ["CellSpecularity", "\t" + " ".join(_fmt(x) for x in m.cellspec)]
    L += ["FluctuationSpecularity", "\t" + _fmt(m.specfluc)]
    L += ["temp # Temperature", "\t" + _fmt(m.temperature)]
    L += ["Np # number of incident phonons", "\t%d" % m.nphonon]
    L += ["Nangle # number of incident angles", "\t%d" % m.nangle]
    L += ["ScopeY", "\t" + _fmt(m.scopey), "ScopeZ", "\t" + _fmt(m.scopez)]
    if m.mfp is not None:
        if len(m.mfp) == 1:
            L += ["MeanFreePath", "\t" + _fmt(m.mfp[0])]
        else:
            L += ["MeanFreePath", "\t%s %s %d %s" % (_fmt(m.mfp[0]), _fmt(m.mfp[1]), m.mfp[2], m.mfp[3])]
    L += ["Trace", "\t%d" % m.trace]
    L += ["", "# materials", "Nm # number of materials", "\t%d" % len(m.materials),
          "material contents"]
    for name in m.materials:
        L += ["\t" + name, "\t\t100.0"]
    L += ["", "# transmission functions", "Nt", "\t%d %d" % m.nt, "materials"]
    for r in m.trans_rows:
        L += ["\t%s %s mode %s Specular" % (r.material1, r.material2, r.kind),
              "\t\t%s %s" % (_fmt(r.value), _fmt(r.specularity))]
    L += ["", "# grains", "Shape # shape of grains", "\t" + m.kind_grain,
          "Ng Ngx Ngy Ngz # number of grains", "\t%d %d 1 1" % (len(m.grains), len(m.grains))]
    for ig, g in enumerate(m.grains, start=1):
        L += ["## grain No.%d" % ig, "  generator", "\t" + " ".join(_fmt(x) for x in g.generatrix)]
        L += ["  " + g.shape, "\t" + " ".join(_fmt(x) for x in g.lengths)]
        L += ["  material", "\t" + g.material]
        if g.shape in ("sphere", "cylinder") and m.kind_grain == "bricks":
            f = g.facets[0]
            L += ["  material_to = %s" % f.material_to, "  grain_to = %d" % f.grain_to]
            continue
        L += ["# facets of grain No.%d" % ig, "  Nf # number of facets", "\t%d" % len(g.facets)]
        for jf, f in enumerate(g.facets, start=1):
            L += ["#  facet No.%d" % jf, "  Nv # number of vertex", "\t\t%d" % len(f.vertices)]
            for kv, v in enumerate(f.vertices, start=1):
                L += ["\tvertex %d" % kv, "\t\t" + " ".join(_fmt(x) for x in v)]
            L += ["  material_to = %s" % f.material_to, "  grain_to = %d" % f.grain_to]
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


class McpError(ValueError):
    pass


def _find_key(lines: List[str], key: str) -> Optional[int]:
    """findKey (RT:212-231): index of the line after the first line whose first token is key."""
    for i, ln in enumerate(lines):
        toks = ln.split()
        if toks and toks[0] == key:
            return i + 1
    return None


def _nums(line: str, conv=float) -> list:
    out = []
    for tok in line.replace(",", " ").split():
        try:
            out.append(conv(tok))
        except ValueError:
            break
    return out


def read_mcp(path: str) -> McpModel:
    with open(path) as fh:
        lines = fh.read().splitlines()
    m = McpModel()

    def value_line(key: str, required: bool = True) -> Optional[str]:
        i = _find_key(lines, key)
        if i is None:
            if required:
                raise McpError("Parameter Error : not found %s" % key)
            return None
        return lines[i]

    m.ncell = tuple(_nums(value_line(k), int)[0] for k in ("NXcell", "NYcell", "NZcell"))
    m.lcell = tuple(_nums(value_line(k))[0] for k in ("LXcell", "LYcell", "LZcell"))
    cs = value_line("CellSpecularity", False)
    if cs is not None:
        v = _nums(cs)
        m.cellspec = tuple(v[:6]) if len(v) >= 6 else (v[0],) * 6
    sf = value_line("FluctuationSpecularity", False)
    m.specfluc = _nums(sf)[0] if sf is not None else 0.0
    m.temperature = _nums(value_line("temp"))[0]
    m.nphonon = _nums(value_line("Np"), int)[0]
    na = value_line("Nangle", False)
    m.nangle = _nums(na, int)[0] if na is not None else 91
    sy, sz = value_line("ScopeY", False), value_line("ScopeZ", False)
    m.scopey = _nums(sy)[0] if sy is not None else 1.0
    m.scopez = _nums(sz)[0] if sz is not None else 1.0
    mf = value_line("MeanFreePath", False)
    if mf is not None:
        toks = mf.split()
        if len(toks) >= 4 and toks[3] in ("linear", "log"):
            m.mfp = (float(toks[0]), float(toks[1]), int(toks[2]), toks[3])
        else:
            m.mfp = (float(toks[0]),)
    tr = value_line("Trace", False)
    m.trace = _nums(tr, int)[0] if tr is not None else 1000

    i = _find_key(lines, "Nm")
    if i is None:
        raise McpError("Parameter Error : not found Nm")
    nm = _nums(lines[i], int)[0]
    i += 2                                   # value line, "material contents"
    m.materials = []
    for _ in range(nm):
        m.materials.append(lines[i].split()[0])
        i += 2                               # name line, skipped value line

    i = _find_key(lines, "Nt")
    if i is None:
        raise McpError("Parameter Error : not found Nt")
    nt = _nums(lines[i], int)
    m.nt = (nt[0], nt[1])
    i += 2                                   # value line, "materials"
    m.trans_rows = []
    for _ in range(nt[0] * nt[1]):
        t = lines[i].split()
        v = _nums(lines[i + 1])
        m.trans_rows.append(TransRow(t[0], t[1], t[3], v[0], v[1]))
        i += 2

    shape = value_line("Shape").split()[0]
    if shape not in ("voronoi", "bricks"):
        raise McpError("Parameter Error : unknown grain Shape: %s" % shape)
    m.kind_grain = shape
    i = _find_key(lines, "Ng")
    ng = _nums(lines[i], int)[0]
    i += 1
    m.grains = []
    for _ in range(ng):
        i += 2                               # "## grain No.*", "generator"
        gen = tuple(_nums(lines[i])[:3]); i += 1
        word = lines[i].split()[0]; i += 1
        lengths = tuple(_nums(lines[i])); i += 1
        i += 1                               # "material"
        material = lines[i].split()[0]; i += 1
        facets = []
        polygon = shape == "voronoi" or word in ("box", "length", "polyhedron")
        if polygon:
            i += 2                           # "# facets ...", "Nf ..."
            nf = _nums(lines[i], int)[0]; i += 1
            for _ in range(nf):
                i += 2                       # "# facet No.*", "Nv ..."
                nv = _nums(lines[i], int)[0]; i += 1
                verts = []
                for _ in range(nv):
                    i += 1                   # "vertex k"
                    verts.append(tuple(_nums(lines[i])[:3])); i += 1
                mat_to = lines[i].split()[2]; i += 1
                g_to = int(lines[i].split()[2]); i += 1
                facets.append(FacetSpec(verts, mat_to, g_to))
        else:
            mat_to = lines[i].split()[2]; i += 1
            g_to = int(lines[i].split()[2]); i += 1
            facets.append(FacetSpec([], mat_to, g_to))
        m.grains.append(GrainSpec(gen, word, lengths, material, facets))
    return m


def trans_kind_id(kind: str) -> int:
    return {"Value": TRANS_VALUE, "Gamma": TRANS_GAMMA, "TBC": TRANS_TBC}[kind]
