"""ctypes binding of libptrans.so (the C-ABI of include/ptrans.h).

The library is built in place (p-trans_amd/libptrans.so, see csrc/Makefile).  There is
no fallback: if the library or a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libptrans.so")

# every entry point declared in include/ptrans.h
EXPORTS = (
    "pt_version", "pt_build_info", "pt_last_error", "pt_philox_fill",
    "pt_ctx_create", "pt_ctx_destroy", "pt_ctx_set_seed", "pt_ctx_set_event_limit",
    "pt_ctx_set_op_counting", "pt_ctx_set_tuning", "pt_ctx_set_tail",
    "pt_geom_create", "pt_geom_destroy", "pt_geom_set_specularity", "pt_surface_setup_polygon",
    "pt_modes_create", "pt_modes_destroy",
    "pt_trace_grid", "pt_replay", "pt_trace_log", "pt_alias_build", "pt_trace_alias",
    "pt_trace_alias_noalloc", "pt_replay_alias",
    "pt_adjoint_alias", "pt_trace_alias_grains",
    "pt_batch_create", "pt_batch_destroy", "pt_batch_info", "pt_batch_set_specularity",
    "pt_batch_set_top_levels",
    "pt_batch_trace", "pt_batch_trace_graph",
)


class PtStats(ctypes.Structure):
    _fields_ = [("paths", ctypes.c_int64), ("events", ctypes.c_int64), ("draws", ctypes.c_int64),
                ("lost", ctypes.c_int64), ("start_mismatch", ctypes.c_int64),
                ("skipped", ctypes.c_int64), ("kernel_ms", ctypes.c_double),
                ("reduce_ms", ctypes.c_double), ("fp64_ops", ctypes.c_int64),
                ("philox_blocks", ctypes.c_int64), ("retests", ctypes.c_int64),
                ("fallbacks", ctypes.c_int64), ("bvh_nodes", ctypes.c_int64),
                ("bvh_tests", ctypes.c_int64), ("event_limited", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


RECORD_DTYPE = np.dtype([("end_side", np.int32), ("flags", np.int32), ("events", np.int64),
                         ("draws", np.int64), ("which_grain", np.int32), ("start_grain", np.int32),
                         ("pos", np.float64, 3), ("dir", np.float64, 3)], align=True)


class PtransError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libptrans error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load():
    """Load libptrans.so and declare the signatures.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libptrans.so is not built (run __graft_entry__.build() or "
                           "make -C p-trans_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, u64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
    L.pt_version.restype = i32
    L.pt_build_info.restype = ctypes.c_char_p
    L.pt_last_error.restype = ctypes.c_size_t
    L.pt_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.pt_ctx_create.argtypes = [ctypes.c_int, u64, ctypes.POINTER(vp)]
    L.pt_ctx_destroy.argtypes = [vp]
    L.pt_ctx_set_seed.argtypes = [vp, u64]
    L.pt_ctx_set_event_limit.argtypes = [vp, i64]
    L.pt_ctx_set_op_counting.argtypes = [vp, i32]
    L.pt_ctx_set_tuning.argtypes = [vp, i32, i32, i32]
    L.pt_ctx_set_tail.argtypes = [vp, i64, i64, i32]
    L.pt_geom_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
    L.pt_geom_destroy.argtypes = [vp]
    L.pt_surface_setup_polygon.argtypes = [i32, vp, vp, vp, vp]
    L.pt_modes_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
    L.pt_modes_destroy.argtypes = [vp]
    L.pt_trace_grid.argtypes = [vp, vp, vp, dbl, i32, vp, i64, i32, vp, vp]
    L.pt_replay.argtypes = [vp, vp, vp, dbl, i32, vp, i64, i64, i64, i32, vp]
    L.pt_alias_build.argtypes = [i64, vp, vp, vp]
    L.pt_trace_alias.argtypes = [vp, vp, vp, i64, i32, vp, vp, vp, vp]
    L.pt_trace_alias_noalloc.argtypes = [vp, vp, vp, i64, i32, vp, vp, vp, vp]
    L.pt_replay_alias.argtypes = [vp, vp, vp, vp, i64, i64, i32, vp]
    L.pt_geom_set_specularity.argtypes = [vp, vp]
    L.pt_adjoint_alias.argtypes = [vp, vp, vp, i64, i32, vp, vp, i32, dbl, i32, vp, vp, vp, vp]
    L.pt_trace_alias_grains.argtypes = [vp, vp, vp, i64, i32, vp, vp, vp, vp, vp]
    L.pt_batch_create.argtypes = [vp, i32, vp, vp, i32, i32, ctypes.POINTER(vp)]
    L.pt_batch_destroy.argtypes = [vp]
    L.pt_batch_info.argtypes = [vp, vp]
    L.pt_batch_set_specularity.argtypes = [vp, vp]
    L.pt_batch_set_top_levels.argtypes = [vp, i32]
    L.pt_batch_trace.argtypes = [vp, vp, vp, i32, vp, vp, vp, vp]
    L.pt_batch_trace_graph.argtypes = [vp, vp, vp, i32, vp, vp, vp, vp]
    L.pt_philox_fill.argtypes = [u64, u64, u64, i64, vp]
    L.pt_trace_log.argtypes = [vp, vp, vp, dbl, i32, vp, i64, i64, i64, i64, i32, i64, vp, vp, vp]
    for name in EXPORTS:
        f = getattr(L, name)
        if f.restype is ctypes.c_int:   # default
            f.restype = ctypes.c_int
    _lib = L
    return L


def last_error() -> str:
    L = load()
    buf = ctypes.create_string_buffer(1024)
    L.pt_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int):
    if rc != 0:
        raise PtransError(rc, last_error())
