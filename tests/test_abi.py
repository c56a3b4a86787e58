"""The C-ABI library: it loads, exports every entry point include/ptrans.h declares,
and its host-only pieces behave (no GPU compute here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from _golden import ROOT
from ptrans import _lib, geometry
from ptrans.geometry import SURFACE_DTYPE, PtGeomDesc, PtModesDesc


def declared_functions():
    with open(os.path.join(ROOT, "include", "ptrans.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|size_t|const char \*)\s*\*?\s*(pt_\w+)\s*\(",
                                 text, re.M)))


def test_library_exports_every_declared_function():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert L.pt_version() == 1


def test_struct_layouts_match_header():
    # sizes printed by sizeof() on include/ptrans.h (gcc, x86-64)
    assert SURFACE_DTYPE.itemsize == 184
    assert ctypes.sizeof(PtGeomDesc) == 112
    assert ctypes.sizeof(PtModesDesc) == 56
    assert _lib.RECORD_DTYPE.itemsize == 80
    assert ctypes.sizeof(_lib.PtStats) == 120


def test_setup_polygon_helper_matches_python_host():
    """pt_surface_setup_polygon (host helper) and the Python host compute the same bits."""
    L = _lib.load()
    rng = np.random.default_rng(3)
    for nv in (3, 4, 5, 7):
        for _ in range(20):
            ang = np.sort(rng.uniform(0, 2 * np.pi, nv))
            v = np.stack([np.cos(ang) * 30, np.sin(ang) * 20, 0.3 * np.cos(ang)], 1) + rng.normal(size=3)
            v = np.ascontiguousarray(v)
            c, n, t = np.zeros(3), np.zeros(3), np.zeros(9)
            assert L.pt_surface_setup_polygon(nv, v.ctypes.data, c.ctypes.data, n.ctypes.data,
                                              t.ctypes.data) == 0
            pc, pn, pt = geometry.setup_polygon([tuple(x) for x in v])
            assert tuple(c) == pc and tuple(n) == pn
            assert tuple(t) == tuple(x for col in pt for x in col)


def test_errors_are_reported_not_printed():
    L = _lib.load()
    assert L.pt_surface_setup_polygon(2, None, None, None, None) == -1
    assert "3 vertices" in _lib.last_error()
    h = ctypes.c_void_p()
    rc = L.pt_ctx_create(-1, 0, ctypes.byref(h))
    assert rc in (-1, -6) and not h.value
