"""Writes the golden vectors under tests/golden/ from the REFERENCE itself.

Runs the replay harness (oracle/replay: the unmodified src/RTMCphon.f90 compiled with
flang, linked with this project's RANDOM_NUMBER interposer) on seeded inputs and stores
what the reference computed:

  <case>/input.mcp, <case>/Si.anphon   the exact inputs the reference read
  <case>/ref_own.npz                   reference linked with pt_math.h transcendentals
  <case>/ref_libm.npz                  reference with the C library's transcendentals
     path, trans, draws                per path (calcTransmit with Np = 1)
     traj_off, traj                    every trace.out position of each path (REAL(4))
     site_off, site                    the source line of every RANDOM_NUMBER call
     grid                              TransmissionFunction counts (nb, nk, nangle)
  geom_cases.txt, geom_ref.txt         reference geometry routines on random cases

Only runs where /root/reference and oracle/_ref exist (the development container):
    make -C oracle ref && python tests/golden/make_golden.py [laws | examples | case NAME...]

  examples/<name>.json                 loadMCP of the reference's own Examples/01-05: the
                                       parameters and every surface record (numbers only)
"""
from __future__ import annotations

import dataclasses
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "p-trans_amd"))

from ptrans import anphon, gen, mcp  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")

def _two_materials():
    """Bricks with Ge spheres and Si z-cylinders in Si (Nm = 2): the reference then keeps
    no phonon table (RT:637-646) and its Gamma law draws an extra number per decision,
    t = 1/(gamma + 1) + u (1 - 1/(gamma + 1)) (RT:1423-1425, 2396-2402)."""
    from ptrans.mcp import TransRow
    m = gen.inclusions(seed=3, nsphere=4, ncyl=2, kind="Value", value=0.3, p=0.5, nphonon=60,
                       nangle=5, mfp=(500.0,))
    grains = [dataclasses.replace(g, material="Ge") if g.shape == "sphere" else g for g in m.grains]
    rows = [TransRow("Si", "Si", "Value", 0.3, 0.5), TransRow("Si", "Ge", "Gamma", 1.5, 0.4),
            TransRow("Ge", "Si", "Gamma", 0.5, 0.6), TransRow("Ge", "Ge", "Value", 0.0, 0.0)]
    return dataclasses.replace(m, materials=["Si", "Ge"], nt=(2, 2), trans_rows=rows, grains=grains)


# name -> (model factory, (nbranch, nkpoint, table seed), mfp_nm, paths seed, npaths, grid seed)
CASES = {
    "c1_wire_ballistic": (lambda: gen.nanowire(p_wall=0.0, nphonon=40, nangle=5),
                          (2, 3, 11), 1.0e10, 1001, 400, 2001),
    "c1_wire_gray": (lambda: gen.nanowire(p_wall=0.3, nphonon=30, nangle=5, mfp=(200.0,)),
                     (2, 3, 12), 200.0, 1002, 300, 2002),
    "c2_film_holes": (lambda: gen.film_holes(seed=2, nphonon=30, nangle=5),
                      (2, 3, 13), 1.0e10, 1003, 300, 2003),
    "c2_film_holes_gray": (lambda: gen.film_holes(seed=2, nphonon=20, nangle=5, p_hole=0.6,
                                                  mfp=(300.0,)),
                           (2, 3, 14), 300.0, 1004, 240, 2004),
    "c3_inclusions_gamma": (lambda: gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.4,
                                                   nphonon=20, nangle=5, mfp=(500.0,)),
                            (2, 3, 15), 500.0, 1005, 240, 2005),
    "c4_voronoi_gamma": (lambda: gen.voronoi_cube(seed=4, kind="Gamma", value=1.5, p=0.2,
                                                  nphonon=20, nangle=5),
                         (2, 3, 16), 1.0e10, 1006, 240, 2006),
    "c4_voronoi_gray": (lambda: gen.voronoi_cube(seed=7, kind="Value", value=0.6, p=0.3,
                                                 nphonon=20, nangle=5, mfp=(400.0,)),
                        (2, 3, 17), 400.0, 1007, 240, 2007),
    "c5_tri_brick": (lambda: gen.tri_brick(nsub=3, jitter=0.0, kind="Value", value=0.3, p=0.5,
                                           nphonon=20, nangle=5, mfp=(600.0,)),
                     (2, 3, 18), 600.0, 1008, 240, 2008),
    # TBC row (GammaDeterm) and FluctuationSpecularity: the set-up draws use the case's
    # seed, so paths and grid share it
    "c4_voronoi_tbc_fluct": (lambda: dataclasses.replace(
                                 gen.voronoi_cube(seed=4, kind="TBC", value=300.0, p=0.4,
                                                  nphonon=20, nangle=5, mfp=(400.0,)),
                                 specfluc=0.1),
                             (2, 3, 19), 400.0, 1009, 240, 1009),
    # two materials: a 1 x 1 table (the reference loads none for Nm > 1)
    "c3_two_materials_gamma": (_two_materials, (1, 1, 20), 500.0, 1010, 240, 2010),
}


def _parse_trace(path):
    blocks, cur = [], []
    with open(path) as f:
        lines = f.read().split("\n")
    blank = 0
    for ln in lines:
        if ln.strip() == "":
            blank += 1
            if blank == 2 and cur:
                blocks.append(cur)
                cur = []
            continue
        blank = 0
        cur.append([float(t) for t in ln.split()])
    if cur:
        blocks.append(cur)
    return blocks


def _parse_sites(path):
    per = []
    with open(path) as f:
        for ln in f:
            ln = ln.strip()
            if ln.startswith("P "):
                per.append([])
            elif ln:
                per[-1].append(int(ln))
    return per


TRACE_TEXT_CASES = ("c1_wire_gray", "c4_voronoi_gamma")


def run_case(name, spec, exe_tag):
    factory, (nb, nk, tseed), mfp, pseed, npaths, gseed = spec
    model = factory()
    outdir = os.path.join(HERE, name)
    os.makedirs(outdir, exist_ok=True)
    mcp.write_mcp(os.path.join(outdir, "input.mcp"), model)
    anphon.write_anphon(os.path.join(outdir, "Si.anphon"), anphon.synthetic_table(nb, nk, tseed))
    exe = os.path.join(REF_DIR, "replay_" + exe_tag)
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "Materials"))
        shutil.copy(os.path.join(outdir, "input.mcp"), os.path.join(tmp, "input.mcp"))
        shutil.copy(os.path.join(outdir, "Si.anphon"), os.path.join(tmp, "Materials", "Si.anphon"))
        env = dict(os.environ, OMP_NUM_THREADS="1")
        out = subprocess.run([exe, "paths", "input.mcp", repr(mfp), str(pseed), "0", str(npaths)],
                             cwd=tmp, env=env, capture_output=True, text=True, timeout=600, check=True)
        recs = [ln.split() for ln in out.stdout.splitlines() if ln.startswith("P ")]
        path = np.array([int(r[1]) for r in recs], dtype=np.int64)
        trans = np.array([int(r[2]) for r in recs], dtype=np.int8)
        draws = np.array([int(r[3]) for r in recs], dtype=np.int64)
        blocks = _parse_trace(os.path.join(tmp, "trace.out"))
        if exe_tag == "own" and name in TRACE_TEXT_CASES:
            # the reference's own trace.out text of the first 40 paths (writer fixture)
            with open(os.path.join(tmp, "trace.out")) as f:
                text = f.read()
            parts = text.split("\n\n\n")
            with open(os.path.join(outdir, "trace_ref.out"), "w") as f:
                f.write("\n\n\n".join(parts[:40]) + "\n\n\n")
        sites = _parse_sites(os.path.join(tmp, "draws.log"))
        if len(blocks) != len(path) or len(sites) != len(path):
            raise RuntimeError("%s: %d paths, %d trace blocks, %d site lists"
                               % (name, len(path), len(blocks), len(sites)))
        for i, s in enumerate(sites):
            if len(s) != draws[i]:
                raise RuntimeError("%s: draw count mismatch on path %d" % (name, i))
        traj_off = np.cumsum([0] + [len(b) for b in blocks]).astype(np.int64)
        traj = np.array([p for b in blocks for p in b], dtype=np.float32).reshape(-1, 3)
        site_off = np.cumsum([0] + [len(s) for s in sites]).astype(np.int64)
        site = np.array([x for s in sites for x in s], dtype=np.int16)
        out = subprocess.run([exe, "grid", "input.mcp", repr(mfp), str(gseed)], cwd=tmp, env=env,
                             capture_output=True, text=True, timeout=600, check=True)
        grid = np.zeros((nb, nk, model.nangle), dtype=np.int64)
        for ln in out.stdout.splitlines():
            if ln.startswith("C "):
                _, ib, ik, ia, c = ln.split()
                grid[int(ib) - 1, int(ik) - 1, int(ia) - 1] = int(c)
    np.savez_compressed(os.path.join(outdir, "ref_%s.npz" % exe_tag), path=path, trans=trans,
                        draws=draws, traj_off=traj_off, traj=traj, site_off=site_off, site=site,
                        grid=grid)
    meta = dict(mfp=mfp, paths_seed=pseed, npaths=npaths, grid_seed=gseed, nbranch=nb,
                nkpoint=nk, nangle=model.nangle, nphonon=model.nphonon)
    with open(os.path.join(outdir, "meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    return int(trans.sum()), len(trans), int(grid.sum())


def geom_cases(seed=77, n=400):
    """Random rays against random polygons, spheres and z-cylinders, plus grazing and
    axis-parallel rays, as input for the reference's own geometry routines."""
    rng = np.random.default_rng(seed)
    L = []

    def f(v):
        return " ".join("%.17g" % x for x in np.ravel(v))

    def unit(v):
        v = np.asarray(v, dtype=float)
        return v / np.linalg.norm(v)

    for i in range(n):
        kind = i % 3
        if kind == 0:
            nv = int(rng.integers(3, 7))
            c = rng.uniform(-50, 50, 3)
            nrm = unit(rng.normal(size=3))
            if i % 30 == 0:
                nrm = np.array([0.0, 0.0, 1.0 if i % 60 == 0 else -1.0])
            u = unit(np.cross(nrm, [1.0, 0.3, 0.1]))
            w = np.cross(nrm, u)
            ang = np.sort(rng.uniform(0, 2 * np.pi, nv))
            rad = rng.uniform(10, 40, nv)
            verts = [c + rad[k] * (np.cos(ang[k]) * u + np.sin(ang[k]) * w) for k in range(nv)]
            rs = c - nrm * rng.uniform(1, 80) + rng.normal(scale=20, size=3)
            nd = unit(nrm + rng.normal(scale=0.8, size=3))
            p = c + rng.normal(scale=25, size=3)
            L += ["POLY", str(nv), f(verts), f(np.concatenate([rs, nd, p]))]
        else:
            c = rng.uniform(-50, 50, 3)
            r = float(rng.uniform(5, 60))
            rs = c + rng.normal(scale=60, size=3)
            if i % 7 == 0:              # start on the surface, moving in or out
                rs = c + r * unit(rng.normal(size=3))
            nd = unit(c - rs + rng.normal(scale=30, size=3))
            if i % 11 == 0:
                nd = np.array([0.0, 0.0, 1.0])      # axis-parallel
            L += ["SPH" if kind == 1 else "CYL", f(np.concatenate([rs, nd, c, [r, 200.0]]))]
    return "\n".join(L) + "\n"


def run_geom():
    text = geom_cases()
    with open(os.path.join(HERE, "geom_cases.txt"), "w") as fh:
        fh.write(text)
    out = subprocess.run([os.path.join(REF_DIR, "replay_libm"), "geom",
                          os.path.join(HERE, "geom_cases.txt")],
                         capture_output=True, text=True, timeout=120, check=True)
    with open(os.path.join(HERE, "geom_ref.txt"), "w") as fh:
        fh.write(out.stdout)


# post-processing cases: name -> (model, table (nb, nk, seed), np, seed)
POST_CASES = {
    "post_matthiessen": (lambda: gen.nanowire(p_wall=0.0, nphonon=60, nangle=11), (3, 12, 21), 60, 31),
    "post_mfp_log": (lambda: gen.film_holes(seed=2, nphonon=40, nangle=7, mfp=(10.0, 1000.0, 3, "log")),
                     (2, 15, 22), 40, 32),
}


def run_post(name, spec):
    """effectiveMFP + fileoutput of the reference on transmission tables traced by the
    oracle (which is pinned to the reference by the path cases above)."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from ptrans import geometry
    factory, (nb, nk, tseed), np_, seed = spec
    model = factory()
    outdir = os.path.join(HERE, name)
    os.makedirs(outdir, exist_ok=True)
    mcp.write_mcp(os.path.join(outdir, "input.mcp"), model)
    table = anphon.synthetic_table(nb, nk, tseed)
    anphon.write_anphon(os.path.join(outdir, "Si.anphon"), table)
    tab = anphon.read_anphon(os.path.join(outdir, "Si.anphon"), "Si")
    geo = geometry.build_geometry(model)
    gdesc, _ga = geo.desc()
    mdesc, _ma = geometry.modes_desc(tab.omega, tab.vel, tab.mfp, tab.weight)
    blocks = []
    for mfp in model.mean_free_paths():
        counts, _ = O.trace_grid(gdesc, mdesc, mfp, model.nangle, np_, seed)
        blocks.append(counts)
    counts = np.stack(blocks, axis=-1)
    np.save(os.path.join(outdir, "counts.npy"), counts)
    with open(os.path.join(outdir, "transmit.txt"), "w") as f:
        for k in range(counts.shape[-1]):
            t = counts[..., k].astype(np.float64) / float(np_)
            f.write(" ".join("%.17g" % x for x in t.reshape(-1, order="F")) + "\n")
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "Materials"))
        shutil.copy(os.path.join(outdir, "input.mcp"), os.path.join(tmp, "input.mcp"))
        shutil.copy(os.path.join(outdir, "Si.anphon"), os.path.join(tmp, "Materials", "Si.anphon"))
        shutil.copy(os.path.join(outdir, "transmit.txt"), os.path.join(tmp, "transmit.txt"))
        out = subprocess.run([os.path.join(REF_DIR, "replay_libm"), "post", "input.mcp",
                              "transmit.txt"], cwd=tmp, env=dict(os.environ, OMP_NUM_THREADS="1"),
                             capture_output=True, text=True, timeout=120, check=True)
        lines = out.stdout.splitlines()
        k = lines.index("SEMFP")
        vals = []
        for ln in lines[k + 1:]:
            try:
                vals.extend(float(x) for x in ln.split())
            except ValueError:
                break
        semfp = np.array(vals[:nk * nb * counts.shape[-1]]).reshape((nk, nb, counts.shape[-1]), order="F")
        np.save(os.path.join(outdir, "semfp_ref.npy"), semfp)
        for fn in ("thermalproperty.out", "effectiveMFP.out", "phononconductivity.out"):
            shutil.copy(os.path.join(tmp, fn), os.path.join(outdir, fn))
    return semfp


# name -> (model factory, (nbranch, nkpoint, table seed), set-up seeds)
LAW_CASES = {
    "voronoi_tbc300_fluct": (lambda: dataclasses.replace(
        gen.voronoi_cube(seed=4, kind="TBC", value=300.0, p=0.4, nphonon=10, nangle=5),
        specfluc=0.1), (3, 50, 5), (7, 8)),
    "inclusions_tbc50_fluct": (lambda: dataclasses.replace(
        gen.inclusions(seed=3, kind="TBC", value=50.0, p=0.5, nphonon=10, nangle=5),
        specfluc=0.25), (2, 40, 6), (11,)),
    "voronoi_tbc1": (lambda: gen.voronoi_cube(seed=5, kind="TBC", value=1.0, p=0.2, nphonon=10,
                                              nangle=5), (3, 60, 7), (1,)),
    "voronoi_tbc1e4": (lambda: gen.voronoi_cube(seed=5, kind="TBC", value=1.0e4, p=0.2,
                                                nphonon=10, nangle=5), (3, 60, 7), (1,)),
}


def run_laws(name, spec):
    """loadMCP of the reference on .mcp inputs with TBC rows and FluctuationSpecularity:
    the gamma of each material pair and every surface's specularity."""
    factory, (nb, nk, tseed), seeds = spec
    model = factory()
    outdir = os.path.join(HERE, "laws", name)
    os.makedirs(outdir, exist_ok=True)
    mcp.write_mcp(os.path.join(outdir, "input.mcp"), model)
    anphon.write_anphon(os.path.join(outdir, "Si.anphon"), anphon.synthetic_table(nb, nk, tseed))
    for seed in seeds:
        with tempfile.TemporaryDirectory() as tmp:
            os.makedirs(os.path.join(tmp, "Materials"))
            shutil.copy(os.path.join(outdir, "input.mcp"), os.path.join(tmp, "input.mcp"))
            shutil.copy(os.path.join(outdir, "Si.anphon"), os.path.join(tmp, "Materials", "Si.anphon"))
            out = subprocess.run([os.path.join(REF_DIR, "replay_libm"), "laws", "input.mcp", str(seed)],
                                 cwd=tmp, capture_output=True, text=True, timeout=120, check=True)
        lines = [ln for ln in out.stdout.splitlines() if ln.startswith(("T ", "S "))]
        with open(os.path.join(outdir, "ref_seed%d.txt" % seed), "w") as f:
            f.write("\n".join(lines) + "\n")
    return len(lines)


EXAMPLES = "/root/reference/Examples"


def _f(x):
    return float(x.replace("D", "E"))


def run_examples():
    """loadMCP of the reference on its own Examples/01-05 (with a synthetic Si table; the
    examples use the Value law, which reads none of it): every parameter the trace reads
    and every surface record, as tests/golden/examples/<name>.json."""
    outdir = os.path.join(HERE, "examples")
    os.makedirs(outdir, exist_ok=True)
    for fn in sorted(os.listdir(EXAMPLES)):
        if not fn.endswith(".mcp"):
            continue
        with tempfile.TemporaryDirectory() as tmp:
            os.makedirs(os.path.join(tmp, "Materials"))
            anphon.write_anphon(os.path.join(tmp, "Materials", "Si.anphon"),
                                anphon.synthetic_table(2, 3, 11))
            out = subprocess.run([os.path.join(REF_DIR, "replay_libm"), "dump",
                                  os.path.join(EXAMPLES, fn)], cwd=tmp, capture_output=True,
                                 text=True, timeout=120, check=True)
        rec = dict(source="Examples/" + fn, mfp=[], grains=[])
        for ln in out.stdout.splitlines():
            t = ln.split()
            if not t:
                continue
            if t[0] == "PARAM":
                v = t[1:]
                rec["param"] = dict(box=[_f(x) for x in v[0:3]], scope=[_f(x) for x in v[3:5]],
                                    temperature=_f(v[5]), nphonon=int(v[6]), nangle=int(v[7]),
                                    nmfp=int(v[8]), kind_grain=int(v[9]), nmaterial=int(v[10]),
                                    specfluc=_f(v[11]))
            elif t[0] == "MFP":
                rec["mfp"].append(_f(t[2]))
            elif t[0] == "CELLSPEC":
                rec["cellspec"] = [_f(x) for x in t[1:]]
            elif t[0] == "GRAINS":
                rec["grain_range"] = [int(x) for x in t[1:]]
            elif t[0] == "G":
                rec["grains"].append(dict(index=int(t[1]), incident=int(t[2]), nsurface=int(t[3]),
                                          generatrix=[_f(x) for x in t[4:7]], surfaces=[]))
            elif t[0] == "S":
                k = [int(x) for x in t[3:]]
                rec["grains"][-1]["surfaces"].append(dict(
                    kind=k[0], internal=k[1], incident=k[2], grain_inside=k[3], grain_outside=k[4],
                    imaterial1=k[5], imaterial2=k[6], trans_kind=k[7], nvertex=k[8], vertices=[]))
            elif t[0] in ("R", "C", "N", "T", "V"):
                sf = rec["grains"][-1]["surfaces"][-1]
                v = [_f(x) for x in t[1:]]
                if t[0] == "R":
                    sf["specularity"], sf["radius"], sf["law"] = v
                elif t[0] == "V":
                    sf["vertices"].append(v)
                else:
                    sf[{"C": "center", "N": "normal", "T": "tangent"}[t[0]]] = v
        with open(os.path.join(outdir, fn[:-4] + ".json"), "w") as f:
            json.dump(rec, f, indent=0)
        print("%-32s %d grains, %d surfaces" % (fn, len(rec["grains"]),
                                                sum(len(g["surfaces"]) for g in rec["grains"])))


def main():
    if not os.path.exists(os.path.join(REF_DIR, "replay_own")):
        sys.exit("oracle/_ref is not built: make -C oracle ref (needs /root/reference)")
    if len(sys.argv) > 1 and sys.argv[1] == "examples":
        run_examples()
        return
    if len(sys.argv) > 2 and sys.argv[1] == "case":
        for name in sys.argv[2:]:
            for tag in ("own", "libm"):
                t, n, g = run_case(name, CASES[name], tag)
                print("%-22s %-4s paths %4d transmitted %4d grid total %d" % (name, tag, n, t, g))
        return
    for name, spec in LAW_CASES.items():
        print("%-22s %d lines" % (name, run_laws(name, spec)))
    if len(sys.argv) > 1 and sys.argv[1] == "laws":
        return
    for name, spec in POST_CASES.items():
        sem = run_post(name, spec)
        print("%-22s SEMFP mean %.6g" % (name, sem.mean()))
    run_geom()
    for name, spec in CASES.items():
        for tag in ("own", "libm"):
            t, n, g = run_case(name, spec, tag)
            print("%-22s %-4s paths %4d transmitted %4d grid total %d" % (name, tag, n, t, g))


if __name__ == "__main__":
    main()
