"""ctypes access to liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as
the checker of libptrans.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MATH_LIBM = 0
MATH_OWN = 1


class PathRecord(ctypes.Structure):
    _fields_ = [("end_side", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("events", ctypes.c_int64), ("draws", ctypes.c_int64),
                ("which_grain", ctypes.c_int32), ("start_grain", ctypes.c_int32),
                ("pos", ctypes.c_double * 3), ("dir", ctypes.c_double * 3)]


RECORD_DTYPE = np.dtype([("end_side", np.int32), ("flags", np.int32), ("events", np.int64),
                         ("draws", np.int64), ("which_grain", np.int32), ("start_grain", np.int32),
                         ("pos", np.float64, 3), ("dir", np.float64, 3)], align=True)
assert RECORD_DTYPE.itemsize == ctypes.sizeof(PathRecord)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, dbl, i32, u64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int, ctypes.c_uint64
        L.oracle_trace_grid.argtypes = [vp, vp, dbl, i32, vp, i64, u64, i32, i64, i64, vp, vp]
        L.oracle_replay.argtypes = [vp, vp, dbl, i32, vp, i64, u64, i32, i64, i64, vp]
        L.oracle_replay_log.argtypes = [vp, vp, dbl, i32, vp, i64, u64, i32, i64, i64, vp,
                                        vp, i64, vp, vp, i64, vp]
        L.oracle_build_alias.argtypes = [i64, vp, vp, vp]
        L.oracle_trace_alias.argtypes = [vp, vp, vp, vp, vp, vp, u64, i32, i64, i64, vp, vp, vp]
        L.oracle_adjoint_alias.argtypes = [vp, vp, vp, vp, vp, vp, u64, i32, i64, i64, vp, i32,
                                           dbl, i32, vp]
        L.oracle_trace_alias_grains.argtypes = [vp, vp, vp, vp, vp, vp, u64, i32, i64, i64, vp,
                                                vp, vp]
        L.oracle_set_event_limit.argtypes = [i64]
        L.oracle_philox_draw.restype = dbl
        L.oracle_philox_draw.argtypes = [u64, u64, u64]
        L.oracle_pt_log.restype = dbl
        L.oracle_pt_log.argtypes = [dbl]
        L.oracle_pt_sincos.argtypes = [dbl, vp, vp]
        L.oracle_setup_polygon.argtypes = [i32, vp, vp, vp, vp]
        L.oracle_dist_polygon.argtypes = [vp, vp, vp, i32, vp, vp]
        L.oracle_inside_polygon.argtypes = [vp, i32, vp]
        L.oracle_dist_sphere.argtypes = [vp, vp, vp, dbl, vp]
        L.oracle_dist_cylinder.argtypes = [vp, vp, vp, dbl, dbl, vp]
        L.oracle_surface_normal.argtypes = [i32, vp, vp, dbl, vp, vp]
        L.oracle_surface_rotation.argtypes = [i32, vp, vp, vp, vp]
        _lib = L
    return _lib


def set_event_limit(limit: int) -> None:
    """Events after which the oracle stops a path (PT_PATH_EVENT_LIMIT); default 1e8."""
    lib().oracle_set_event_limit(int(limit))


def _p(a):
    return a.ctypes.data if a is not None else None


def trace_grid(gdesc, mdesc, mfp, nangle, np_, seed, math=MATH_OWN, vangle=None,
               cell_begin=0, cell_end=-1):
    """oracle_trace_grid: per-cell transmitted counts (nb, nk, nangle) and stats."""
    nb, nk = mdesc.nbranch, mdesc.nkpoint
    counts = np.zeros(nb * nk * nangle, dtype=np.int64)
    stats = np.zeros(6, dtype=np.int64)
    va = None if vangle is None else np.ascontiguousarray(vangle, dtype=np.float64)
    rc = lib().oracle_trace_grid(ctypes.addressof(gdesc), ctypes.addressof(mdesc), float(mfp),
                                 int(nangle), _p(va), int(np_), int(seed), int(math),
                                 int(cell_begin), int(cell_end), _p(counts), _p(stats))
    if rc != 0:
        raise RuntimeError("oracle_trace_grid failed")
    return counts.reshape((nb, nk, nangle), order="F"), stats


def replay(gdesc, mdesc, mfp, nangle, np_, seed, first, n, math=MATH_OWN, vangle=None):
    out = np.zeros(n, dtype=RECORD_DTYPE)
    va = None if vangle is None else np.ascontiguousarray(vangle, dtype=np.float64)
    rc = lib().oracle_replay(ctypes.addressof(gdesc), ctypes.addressof(mdesc), float(mfp),
                             int(nangle), _p(va), int(np_), int(seed), int(math),
                             int(first), int(n), _p(out))
    if rc != 0:
        raise RuntimeError("oracle_replay failed")
    return out


def replay_log(gdesc, mdesc, mfp, nangle, np_, seed, first, n, math=MATH_OWN, vangle=None,
               traj_cap=200000, site_cap=400000):
    """oracle_replay plus every trace position and draw site of each path."""
    out = np.zeros(n, dtype=RECORD_DTYPE)
    va = None if vangle is None else np.ascontiguousarray(vangle, dtype=np.float64)
    traj = np.zeros((traj_cap, 3), dtype=np.float64)
    toff = np.zeros(n + 1, dtype=np.int64)
    sites = np.zeros(site_cap, dtype=np.int16)
    soff = np.zeros(n + 1, dtype=np.int64)
    rc = lib().oracle_replay_log(ctypes.addressof(gdesc), ctypes.addressof(mdesc), float(mfp),
                                 int(nangle), _p(va), int(np_), int(seed), int(math),
                                 int(first), int(n), _p(out), _p(traj), traj_cap, _p(toff),
                                 _p(sites), site_cap, _p(soff))
    if rc != 0:
        raise RuntimeError("oracle_replay_log failed (%d)" % rc)
    return out, traj[:toff[-1]], toff, sites[:soff[-1]], soff


def build_alias(weights):
    w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
    prob = np.zeros(w.size)
    alias = np.zeros(w.size, dtype=np.int32)
    if lib().oracle_build_alias(w.size, _p(w), _p(prob), _p(alias)) != 0:
        raise RuntimeError("oracle_build_alias failed")
    return prob, alias


def trace_alias(gdesc, mdesc, mfp_nm, prob, alias, seed, first, n, scale=None, math=MATH_OWN,
                records=False):
    """oracle_trace_alias: per-mode (nb, nk) paths and transmitted counts (+ records)."""
    nb, nk = mdesc.nbranch, mdesc.nkpoint
    mfp = np.ascontiguousarray(np.asarray(mfp_nm, dtype=np.float64).reshape(-1, order="F"))
    prob = np.ascontiguousarray(prob, dtype=np.float64)
    alias = np.ascontiguousarray(alias, dtype=np.int32)
    sc = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
    mp = np.zeros(nb * nk, dtype=np.int64)
    mt = np.zeros(nb * nk, dtype=np.int64)
    rec = np.zeros(n, dtype=RECORD_DTYPE) if records else None
    lib().oracle_trace_alias(ctypes.addressof(gdesc), ctypes.addressof(mdesc), _p(mfp), _p(prob),
                             _p(alias), _p(sc), int(seed), int(math), int(first), int(n),
                             _p(mp), _p(mt), _p(rec))
    out = (mp.reshape((nb, nk), order="F"), mt.reshape((nb, nk), order="F"))
    return out + (rec,) if records else out


def adjoint_alias(gdesc, mdesc, mfp_nm, prob, alias, seed, first, n, group=None, ngroup=0,
                  baseline=0.0, nbatch=64, scale=None, math=MATH_OWN):
    """oracle_adjoint_alias: per-batch sums (nbatch, 3 + ngroup + nb) of
    [1, x, x^2, (x - baseline) * score_k], paths in order."""
    nb = mdesc.nbranch
    mfp = np.ascontiguousarray(np.asarray(mfp_nm, dtype=np.float64).reshape(-1, order="F"))
    prob = np.ascontiguousarray(prob, dtype=np.float64)
    alias = np.ascontiguousarray(alias, dtype=np.int32)
    sc = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
    grp = None if not ngroup else np.ascontiguousarray(group, dtype=np.int32)
    sums = np.zeros((nbatch, 3 + ngroup + nb), dtype=np.float64)
    rc = lib().oracle_adjoint_alias(ctypes.addressof(gdesc), ctypes.addressof(mdesc), _p(mfp),
                                    _p(prob), _p(alias), _p(sc), int(seed), int(math), int(first),
                                    int(n), _p(grp), int(ngroup), float(baseline), int(nbatch),
                                    _p(sums))
    if rc != 0:
        raise ValueError("oracle_adjoint_alias: bad arguments")
    return sums


def trace_alias_grains(gdesc, mdesc, mfp_nm, prob, alias, seed, first, n, ngrain_total,
                       scale=None, math=MATH_OWN):
    """oracle_trace_alias_grains: per-mode (nb, nk) paths, transmitted, and per-grain
    (ngrain_total, 2) sums of [dx / mfp_m, l / mfp_m]."""
    nb, nk = mdesc.nbranch, mdesc.nkpoint
    mfp = np.ascontiguousarray(np.asarray(mfp_nm, dtype=np.float64).reshape(-1, order="F"))
    prob = np.ascontiguousarray(prob, dtype=np.float64)
    alias = np.ascontiguousarray(alias, dtype=np.int32)
    sc = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
    mp = np.zeros(nb * nk, dtype=np.int64)
    mt = np.zeros(nb * nk, dtype=np.int64)
    gt = np.zeros((ngrain_total, 2), dtype=np.float64)
    lib().oracle_trace_alias_grains(ctypes.addressof(gdesc), ctypes.addressof(mdesc), _p(mfp),
                                    _p(prob), _p(alias), _p(sc), int(seed), int(math), int(first),
                                    int(n), _p(mp), _p(mt), _p(gt))
    return mp.reshape((nb, nk), order="F"), mt.reshape((nb, nk), order="F"), gt


def philox_draw(seed, path, draw):
    return lib().oracle_philox_draw(seed, path, draw)
