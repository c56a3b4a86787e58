"""Loading of the golden cases written by tests/golden/make_golden.py."""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
for p in (os.path.join(ROOT, "p-trans_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

from ptrans import anphon, geometry, mcp  # noqa: E402

CASES = sorted(d for d in os.listdir(GOLDEN)
               if os.path.isdir(os.path.join(GOLDEN, d)) and os.path.exists(os.path.join(GOLDEN, d, "meta.json")))


class Case:
    def __init__(self, name):
        self.name = name
        d = os.path.join(GOLDEN, name)
        self.dir = d
        with open(os.path.join(d, "meta.json")) as f:
            self.meta = json.load(f)
        self.model = mcp.read_mcp(os.path.join(d, "input.mcp"))
        self.table = anphon.read_anphon(os.path.join(d, "Si.anphon"), "Si")
        # TBC rows need the bulk table; FluctuationSpecularity draws from the set-up stream
        # of the seed the reference ran with (paths and grid share it where it matters)
        self.geom = geometry.build_geometry(self.model, tables={"Si": self.table},
                                            setup_seed=self.meta["paths_seed"])
        self.gdesc, self._garr = self.geom.desc()
        t = self.table
        self.mdesc, self._marr = geometry.modes_desc(t.omega, t.vel, t.mfp, t.weight)

    def ref(self, tag):
        return dict(np.load(os.path.join(self.dir, "ref_%s.npz" % tag)))

    @property
    def mfp(self):
        return self.meta["mfp"]
