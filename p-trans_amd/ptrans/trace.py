"""The trace through libptrans: the host-side mirror of TransmissionFunction /
calcTransmit (src/RTMCphon.f90:1544-1604, 1608-2414).

    with Context(seed) as ctx:
        geom = ctx.geometry(build_geometry(read_mcp("x.mcp")))
        modes = ctx.modes(read_anphon("Materials/Si.anphon"))
        transmit = ctx.transmission_function(geom, modes, mfp_nm=1e10, nangle=11, np_=1000)

transmission_function returns transmit(nbranch, nkpoint, nangle) as the reference's
module array (count/np per cell, last angle 0), and the raw int64 counts.
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np

from . import _lib
from .constants import BALLISTIC_MFP, PREC_FP64
from .geometry import Geometry, modes_desc


def _stats(stats) -> dict:
    """The stats dict of a trace call; warns when paths stopped at the event limit (the
    library counts them as not transmitted, which biases the transmission low)."""
    d = stats.as_dict()
    if d.get("event_limited", 0) > 0:
        warnings.warn("%d paths stopped at the event limit and were counted as not "
                      "transmitted" % d["event_limited"], RuntimeWarning, stacklevel=3)
    return d


class Geom:
    def __init__(self, ctx, geometry: Geometry):
        self.ctx = ctx
        self.geometry = geometry
        self._desc, self._arrays = geometry.desc()
        h = ctypes.c_void_p()
        _lib.check(_lib.load().pt_geom_create(ctx.handle, ctypes.addressof(self._desc), ctypes.byref(h)))
        self.handle = h

    def set_specularity(self, specularity):
        """pt_geom_set_specularity: new surface%specularity for every surface."""
        p = np.ascontiguousarray(specularity, dtype=np.float64).reshape(-1)
        if p.size != len(self._arrays["surfaces"]):
            raise ValueError("need one specularity per surface (%d)" % len(self._arrays["surfaces"]))
        _lib.check(_lib.load().pt_geom_set_specularity(self.handle, p.ctypes.data))
        self._arrays["surfaces"]["specularity"] = p

    @property
    def specularity(self):
        return self._arrays["surfaces"]["specularity"].copy()

    def close(self):
        if self.handle:
            _lib.load().pt_geom_destroy(self.handle)
            self.handle = None


class Modes:
    def __init__(self, ctx, omega, vel=None, mfp=None, weight=None, max_omega=None,
                 temperature=300.0):
        self.ctx = ctx
        omega = np.asarray(omega, dtype=np.float64)
        self.nbranch, self.nkpoint = omega.shape
        self.mfp = None if mfp is None else np.asarray(mfp, dtype=np.float64)
        self._desc, self._arrays = modes_desc(omega, vel, mfp, weight, max_omega, temperature)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().pt_modes_create(ctx.handle, ctypes.addressof(self._desc), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            _lib.load().pt_modes_destroy(self.handle)
            self.handle = None


class Context:
    """One GPU, one stream, one Philox seed (pt_ctx)."""

    def __init__(self, seed: int = 0, device: int = 0):
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.pt_ctx_create(device, seed, ctypes.byref(h)))
        self.handle = h
        self._owned = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        for o in self._owned:
            o.close()
        self._owned = []
        if self.handle:
            _lib.load().pt_ctx_destroy(self.handle)
            self.handle = None

    def set_seed(self, seed: int):
        _lib.check(_lib.load().pt_ctx_set_seed(self.handle, seed))

    def set_event_limit(self, limit: int):
        _lib.check(_lib.load().pt_ctx_set_event_limit(self.handle, limit))

    def set_tuning(self, chunk: int = 256, refill_min: int = 16, blocks_per_cu: int = 0):
        _lib.check(_lib.load().pt_ctx_set_tuning(self.handle, chunk, refill_min, blocks_per_cu))

    def set_tail(self, events: int = 500, max_records: int = 1 << 20, live_blocks: int = 16):
        """Paths past `events` events (and the last few of each wave) go to the tail kernel;
        at most max_records per launch (0: none); live_blocks workgroups of it run beside
        the persistent kernel (0: it runs after it only)."""
        _lib.check(_lib.load().pt_ctx_set_tail(self.handle, events, max_records, live_blocks))

    def set_op_counting(self, on: bool):
        _lib.check(_lib.load().pt_ctx_set_op_counting(self.handle, int(bool(on))))

    def geometry(self, geometry: Geometry) -> Geom:
        g = Geom(self, geometry)
        self._owned.append(g)
        return g

    def modes(self, table=None, omega=None, max_omega=None, vel=None, mfp=None, weight=None,
              temperature=300.0) -> Modes:
        """Upload a bulk table (an Anphon, mfp column in nm) or bare arrays; vel, mfp and
        weight enable the alias-mode trace."""
        if table is not None:
            m = Modes(self, table.omega, table.vel, table.mfp, table.weight, max_omega, temperature)
        else:
            m = Modes(self, omega, vel, mfp, weight, max_omega, temperature)
        self._owned.append(m)
        return m

    def trace_grid(self, geom: Geom, modes: Modes, mfp_nm: float, nangle: int, np_: int,
                   vangle=None, precision: int = PREC_FP64):
        """pt_trace_grid: counts (nb, nk, nangle) int64 and the stats dict."""
        nb, nk = modes.nbranch, modes.nkpoint
        counts = np.zeros(nb * nk * nangle, dtype=np.int64)
        stats = _lib.PtStats()
        va = None
        if vangle is not None:
            va = np.ascontiguousarray(vangle, dtype=np.float64)
        _lib.check(_lib.load().pt_trace_grid(
            self.handle, geom.handle, modes.handle, float(mfp_nm), int(nangle),
            va.ctypes.data if va is not None else None, int(np_), int(precision),
            counts.ctypes.data, ctypes.byref(stats)))
        return counts.reshape((nb, nk, nangle), order="F"), _stats(stats)

    def transmission_function(self, geom: Geom, modes: Modes, mfp_nm: float = BALLISTIC_MFP,
                              nangle: int = 91, np_: int = 1000, vangle=None):
        """TransmissionFunction (RT:1544-1604): transmit(ib, ik, ia) = count/np."""
        counts, stats = self.trace_grid(geom, modes, mfp_nm, nangle, np_, vangle)
        transmit = counts.astype(np.float64) / float(np_) if np_ > 0 else counts * 0.0
        return transmit, counts, stats

    def replay(self, geom: Geom, modes: Modes, mfp_nm: float, nangle: int, np_: int,
               first: int, n: int, vangle=None, precision: int = PREC_FP64):
        """pt_replay: per-path records of paths first .. first+n-1."""
        out = np.zeros(n, dtype=_lib.RECORD_DTYPE)
        va = None
        if vangle is not None:
            va = np.ascontiguousarray(vangle, dtype=np.float64)
        _lib.check(_lib.load().pt_replay(
            self.handle, geom.handle, modes.handle, float(mfp_nm), int(nangle),
            va.ctypes.data if va is not None else None, int(np_), int(first), int(n),
            int(precision), out.ctypes.data))
        return out

    def trace_log(self, geom: Geom, modes: Modes, mfp_nm: float, nangle: int, np_: int,
                  first: int, n: int, stride: int = 1, max_points: int = 4096, vangle=None,
                  precision: int = PREC_FP64):
        """pt_trace_log: the trace.out positions of paths first + i * stride (i < n): a
        list of (npoints_i, 3) FP64 arrays (cut at max_points) and the per-path records."""
        pts = np.zeros((n, max_points, 3), dtype=np.float64)
        npts = np.zeros(n, dtype=np.int64)
        rec = np.zeros(n, dtype=_lib.RECORD_DTYPE)
        va = None
        if vangle is not None:
            va = np.ascontiguousarray(vangle, dtype=np.float64)
        _lib.check(_lib.load().pt_trace_log(
            self.handle, geom.handle, modes.handle, float(mfp_nm), int(nangle),
            va.ctypes.data if va is not None else None, int(np_), int(first), int(stride), int(n),
            int(precision), int(max_points), pts.ctypes.data, npts.ctypes.data, rec.ctypes.data))
        return [pts[i, :min(int(npts[i]), max_points)].copy() for i in range(n)], rec

    def trace_alias(self, geom: Geom, modes: Modes, npaths: int, branch_scale=None,
                    precision: int = PREC_FP64, noalloc: bool = False):
        """pt_trace_alias (noalloc: pt_trace_alias_noalloc, the allocation-free variant):
        per-mode (nb, nk) int64 path and transmitted counts, stats."""
        nb, nk = modes.nbranch, modes.nkpoint
        paths = np.zeros(nb * nk, dtype=np.int64)
        trans = np.zeros(nb * nk, dtype=np.int64)
        stats = _lib.PtStats()
        sc = None if branch_scale is None else np.ascontiguousarray(branch_scale, dtype=np.float64)
        L = _lib.load()
        fn = L.pt_trace_alias_noalloc if noalloc else L.pt_trace_alias
        _lib.check(fn(
            self.handle, geom.handle, modes.handle, int(npaths), int(precision),
            sc.ctypes.data if sc is not None else None, paths.ctypes.data, trans.ctypes.data,
            ctypes.byref(stats)))
        return (paths.reshape((nb, nk), order="F"), trans.reshape((nb, nk), order="F"),
                _stats(stats))

    def replay_alias(self, geom: Geom, modes: Modes, first: int, n: int, branch_scale=None,
                     precision: int = PREC_FP64):
        out = np.zeros(n, dtype=_lib.RECORD_DTYPE)
        sc = None if branch_scale is None else np.ascontiguousarray(branch_scale, dtype=np.float64)
        _lib.check(_lib.load().pt_replay_alias(
            self.handle, geom.handle, modes.handle, sc.ctypes.data if sc is not None else None,
            int(first), int(n), int(precision), out.ctypes.data))
        return out


def _adjoint_alias(self, geom: Geom, modes: Modes, npaths: int, surface_group=None,
                   ngroup: int = 0, baseline: float = 0.0, nbatch: int = 64, branch_scale=None,
                   precision: int = PREC_FP64):
    """pt_adjoint_alias: per-mode tallies, per-batch sums (nbatch, 3 + ngroup + nb) of
    [1, x, x^2, (x - baseline) * score_k] with x = T / mfp_m, and the stats."""
    nb, nk = modes.nbranch, modes.nkpoint
    paths = np.zeros(nb * nk, dtype=np.int64)
    trans = np.zeros(nb * nk, dtype=np.int64)
    sums = np.zeros((nbatch, 3 + ngroup + nb), dtype=np.float64)
    stats = _lib.PtStats()
    sc = None if branch_scale is None else np.ascontiguousarray(branch_scale, dtype=np.float64)
    grp = None
    if ngroup:
        grp = np.ascontiguousarray(surface_group, dtype=np.int32)
    _lib.check(_lib.load().pt_adjoint_alias(
        self.handle, geom.handle, modes.handle, int(npaths), int(precision),
        sc.ctypes.data if sc is not None else None, grp.ctypes.data if grp is not None else None,
        int(ngroup), float(baseline), int(nbatch), paths.ctypes.data, trans.ctypes.data,
        sums.ctypes.data, ctypes.byref(stats)))
    return (paths.reshape((nb, nk), order="F"), trans.reshape((nb, nk), order="F"), sums,
            _stats(stats))


Context.adjoint_alias = _adjoint_alias


def _trace_alias_grains(self, geom: Geom, modes: Modes, npaths: int, branch_scale=None,
                        precision: int = PREC_FP64):
    """pt_trace_alias_grains: per-mode tallies, per-grain (ngrain_total, 2) sums of
    [dx / mfp_m, l / mfp_m] over the flights inside each grain, and the stats."""
    nb, nk = modes.nbranch, modes.nkpoint
    paths = np.zeros(nb * nk, dtype=np.int64)
    trans = np.zeros(nb * nk, dtype=np.int64)
    tally = np.zeros((len(geom.geometry.grains), 2), dtype=np.float64)
    stats = _lib.PtStats()
    sc = None if branch_scale is None else np.ascontiguousarray(branch_scale, dtype=np.float64)
    _lib.check(_lib.load().pt_trace_alias_grains(
        self.handle, geom.handle, modes.handle, int(npaths), int(precision),
        sc.ctypes.data if sc is not None else None, paths.ctypes.data, trans.ctypes.data,
        tally.ctypes.data, ctypes.byref(stats)))
    return (paths.reshape((nb, nk), order="F"), trans.reshape((nb, nk), order="F"), tally,
            _stats(stats))


Context.trace_alias_grains = _trace_alias_grains


class Batch:
    """pt_batch: mesh structures (same box) traced together in one persistent launch, each
    structure owning a contiguous range of alias-mode path ids."""

    def __init__(self, ctx, geometries, npaths, leaf_size: int = 4, grid: int = 64):
        self.ctx = ctx
        self.geometries = list(geometries)
        self._descs = [g.desc() for g in self.geometries]
        n = len(self._descs)
        ptrs = (ctypes.c_void_p * n)(*[ctypes.addressof(d) for d, _ in self._descs])
        self.npaths = np.ascontiguousarray(np.broadcast_to(np.asarray(npaths, dtype=np.int64), (n,)))
        self.path_off = np.concatenate([[0], np.cumsum(self.npaths)])
        h = ctypes.c_void_p()
        _lib.check(_lib.load().pt_batch_create(ctx.handle, n, ctypes.addressof(ptrs),
                                                self.npaths.ctypes.data, int(leaf_size), int(grid),
                                                ctypes.byref(h)))
        self.handle = h
        self.nstruct = n
        self.leaf_size, self.grid = int(leaf_size), int(grid)
        ctx._owned.insert(0, self)          # closed before the geometries and the context

    def info(self):
        v = np.zeros(6, dtype=np.int64)
        _lib.check(_lib.load().pt_batch_info(self.handle, v.ctypes.data))
        return dict(zip(("surfaces", "grains", "nodes", "prims", "depth", "paths"), v.tolist()))

    def set_specularity(self, mean):
        m = np.ascontiguousarray(mean, dtype=np.float64)
        _lib.check(_lib.load().pt_batch_set_specularity(self.handle, m.ctypes.data))

    def set_top_levels(self, levels: int):
        """Top BVH levels each workgroup stages in LDS for the structure it traces (0: none,
        at most 3).  Results do not depend on it."""
        _lib.check(_lib.load().pt_batch_set_top_levels(self.handle, int(levels)))

    def trace(self, modes: Modes, precision: int = PREC_FP64, baseline=None, graph: bool = False):
        """One launch: per-structure sums (nstruct, 5) of [paths, sum x, sum x^2,
        sum (x - baseline) score, sum ((x - baseline) score)^2], completion times (ms) and
        the stats."""
        sums = np.zeros((self.nstruct, 5), dtype=np.float64)
        fin = np.zeros(self.nstruct, dtype=np.float64)
        st = _lib.PtStats()
        b = None if baseline is None else np.ascontiguousarray(baseline, dtype=np.float64)
        fn = _lib.load().pt_batch_trace_graph if graph else _lib.load().pt_batch_trace
        _lib.check(fn(self.ctx.handle, self.handle, modes.handle, int(precision),
                      b.ctypes.data if b is not None else None, sums.ctypes.data, fin.ctypes.data,
                      ctypes.byref(st)))
        return sums, fin, _stats(st)

    def close(self):
        if self.handle:
            _lib.load().pt_batch_destroy(self.handle)
            self.handle = None


def alias_table(weights):
    """pt_alias_build on a flat weight vector: (prob, alias)."""
    w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
    prob = np.zeros(w.size)
    alias = np.zeros(w.size, dtype=np.int32)
    _lib.check(_lib.load().pt_alias_build(w.size, w.ctypes.data, prob.ctypes.data, alias.ctypes.data))
    return prob, alias


def alias_probabilities(prob, alias):
    """The distribution a Vose table samples: (prob[i] + sum_{alias[j]=i} (1-prob[j])) / n."""
    n = prob.size
    p = prob.astype(np.float64).copy()
    np.add.at(p, alias, 1.0 - prob)          # an entry with prob = 1 gives nothing away
    return p / n
