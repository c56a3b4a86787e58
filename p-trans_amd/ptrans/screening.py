"""Structural screening with kappa gradients.

For each candidate structure (a C3-type cube of spheres and z-cylinders, one seed per
structure), kappa/kappa_bulk is differentiated with respect to the specularity of every
inclusion by the adjoint trace (KappaModel), and a few projected Adam steps move the
specularities toward the lowest (or highest) kappa.  Structures are then ranked by the
kappa they reach.

    python -m ptrans.screening --structures 8 --steps 4 --paths 1000000

prints one JSON line per structure and a final ranking line.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from . import gen, geometry
from .anphon import synthetic_table
from .autograd import KappaModel
from .constants import PREC_FP32_RETEST
from .trace import Context

M_TWOPI = 3.14159265358979323846 * 2.0


def screen(structures: int = 8, steps: int = 4, npaths: int = 1_000_000, lr: float = 0.15,
           minimize: bool = True, nbranch: int = 6, nkpoint: int = 200, seed: int = 1,
           device: int = 0, precision: int = PREC_FP32_RETEST, p_bounds=(0.02, 0.98),
           out=None):
    """Returns a list of per-structure dicts sorted by final kappa (ascending when
    minimizing).  Every evaluation uses common random numbers within a structure."""
    t = synthetic_table(nbranch, nkpoint, seed)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    results = []
    with Context(seed=seed, device=device) as ctx:
        modes = ctx.modes(omega=omega, vel=t.vel, mfp=t.mfp_nm, weight=t.weight)
        for j in range(structures):
            model = gen.inclusions(seed=100 + j, kind="Gamma", value=2.0, p=0.5, nphonon=1,
                                   nangle=3)
            geo = geometry.build_geometry(model)
            geom = ctx.geometry(geo)
            group, ngroup = geometry.inclusion_groups(geo)
            km = KappaModel(ctx, geom, modes, group, ngroup, npaths=npaths, precision=precision,
                            seed=seed + 1000 * j)
            p = torch.full((ngroup,), 0.5, dtype=torch.float64, requires_grad=True)
            s = torch.ones(nbranch, dtype=torch.float64)
            opt = torch.optim.Adam([p], lr=lr)
            hist = []
            t0 = time.time()
            for it in range(steps + 1):
                opt.zero_grad()
                k = km(p, s)
                hist.append((float(k.detach()), km.last.stderr))
                if it == steps:
                    break
                (k if minimize else -k).backward()
                opt.step()
                with torch.no_grad():
                    p.clamp_(*p_bounds)
            rec = dict(structure=j, seed=100 + j, inclusions=ngroup, kappa_ratio_start=hist[0][0],
                       kappa_ratio=hist[-1][0], stderr=hist[-1][1],
                       specularity=[round(float(v), 4) for v in p.detach()],
                       grad_norm=float(np.linalg.norm(km.last.grad_p)),
                       evaluations=steps + 1, paths_per_eval=npaths,
                       seconds=round(time.time() - t0, 3))
            results.append(rec)
            if out is not None:
                print(json.dumps(rec), file=out, flush=True)
            geom.close()
    results.sort(key=lambda r: r["kappa_ratio"], reverse=not minimize)
    return results


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--structures", type=int, default=8)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--paths", type=float, default=1e6)
    ap.add_argument("--lr", type=float, default=0.15)
    ap.add_argument("--maximize", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    res = screen(a.structures, a.steps, int(a.paths), a.lr, not a.maximize, seed=a.seed,
                 out=sys.stdout)
    print(json.dumps({"ranking": [r["structure"] for r in res],
                      "best_kappa_ratio": res[0]["kappa_ratio"] if res else math.nan}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
