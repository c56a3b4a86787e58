"""Launch and BVH tuning of the batched trace, by measurement.

tune_batch() times the batch trace of the actual structures over a grid of BVH leaf sizes
(the BVHs are rebuilt), top BVH levels staged in LDS, and persistent-kernel tiles (paths a
wave takes from the queue, idle lanes that trigger a refill, workgroups per CU) on a short
probe run, and returns
the fastest setting with the whole table.  Every setting gives the same results (the
scan result and the path streams do not depend on them), so the choice is by speed
alone.  Results can be cached in a JSON file keyed by the batch's shape.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from .constants import PREC_FP32_RETEST


def _key(geometries, precision):
    h = hashlib.sha1()
    for g in geometries:
        a = g.flatten()
        h.update(np.int64(len(a["surfaces"])).tobytes())
        h.update(np.int64(len(a["grain_offset"])).tobytes())
    h.update(np.int64(precision).tobytes())
    return h.hexdigest()


def tune_batch(ctx, geometries, modes, probe_paths: int = 20000, leaf_sizes=(1, 2, 4, 8),
               chunks=(64, 256, 1024), refills=(8, 16, 32), blocks_per_cu=(0,), top_levels=(0, 1),
               precision: int = PREC_FP32_RETEST, repeats: int = 2, grid: int = 64,
               cache: str | None = None, log=None):
    """Returns dict(leaf_size, top_levels, chunk, refill_min, blocks_per_cu, paths_per_s,
    table)."""
    from .trace import Batch
    key = _key(geometries, precision)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            got = json.load(f).get(key)
        if got:
            return got
    table = []
    best = None
    for leaf in leaf_sizes:
        b = Batch(ctx, geometries, probe_paths, leaf_size=leaf, grid=grid)
        for levels in top_levels:
            b.set_top_levels(levels)
            b.trace(modes, precision=precision, graph=True)              # warm-up, graph
            for chunk in chunks:
                for refill in refills:
                    for bpc in blocks_per_cu:
                        ctx.set_tuning(chunk, refill, bpc)
                        ms = []
                        for _ in range(repeats):
                            t0 = time.perf_counter()
                            _, _, st = b.trace(modes, precision=precision, graph=True)
                            ms.append(min(st["kernel_ms"], (time.perf_counter() - t0) * 1e3))
                        rate = b.npaths.sum() / (min(ms) * 1e-3)
                        row = dict(leaf_size=leaf, top_levels=levels, chunk=chunk,
                                   refill_min=refill, blocks_per_cu=bpc, kernel_ms=min(ms),
                                   paths_per_s=rate)
                        table.append(row)
                        if log:
                            log(row)
                        if best is None or rate > best["paths_per_s"]:
                            best = dict(row)
        b.close()
    ctx.set_tuning(best["chunk"], best["refill_min"], best["blocks_per_cu"])
    best["table"] = table
    if cache:
        data = {}
        if os.path.exists(cache):
            with open(cache) as f:
                data = json.load(f)
        data[key] = best
        with open(cache, "w") as f:
            json.dump(data, f, indent=1)
    return best
