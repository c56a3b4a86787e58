"""Edge cases of the C-ABI: empty and ragged inputs, the size limits, invalid arguments
and handles of another context.  Each failure is an error code with a message (nothing
is launched), each empty call returns zeros."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from ptrans import _lib, gen, geometry
from test_batch import _lattices, _modes, _oracle_struct_sums

OM = np.linspace(2e12, 9e13, 6).reshape(2, 3)


def _film():
    return geometry.build_geometry(gen.film_holes(seed=2, nphonon=10, nangle=5))


@pytest.mark.gpu
def test_gpu_empty_calls_return_zeros(gpu_ctx):
    geo = _film()
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=OM)
    counts, st = gpu_ctx.trace_grid(g, m, 300.0, 5, 0)
    assert counts.shape == (2, 3, 5) and counts.sum() == 0 and st["paths"] == 0
    assert len(gpu_ctx.replay(g, m, 300.0, 5, 10, 0, 0)) == 0
    tr, rec = gpu_ctx.trace_log(g, m, 300.0, 5, 10, 0, 0)
    assert tr == [] and len(rec) == 0
    omega, vel, mfp, weight, _, _ = _modes()
    ma = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    mp, mt, st = gpu_ctx.trace_alias(g, ma, 0)
    assert mp.sum() == 0 and mt.sum() == 0 and st["paths"] == 0
    _, _, sums, st = gpu_ctx.adjoint_alias(g, ma, 0)
    assert not np.any(sums) and st["paths"] == 0


@pytest.mark.gpu
def test_gpu_ragged_batch_with_empty_structures(gpu_ctx):
    """Structures with 0, 57 and 1 paths (and a different mesh each): per-structure sums
    equal the oracle's, empty structures stay zero with no completion time."""
    from ptrans.trace import Batch
    geos = _lattices()
    mods = _modes()
    omega, vel, mfp, weight, _, _ = mods
    npaths = [0, 57, 1]
    gpu_ctx.set_seed(5)
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    b = Batch(gpu_ctx, geos, npaths, leaf_size=2, grid=8)
    for graph in (False, True):
        sums, fin, st = b.trace(m, precision=1, graph=graph)
        assert st["paths"] == 58
        assert not np.any(sums[0]) and fin[0] == -1.0
        for s in (1, 2):
            n, sx, _ = _oracle_struct_sums(geos[s], mods, 5, b.path_off[s], npaths[s])
            assert sums[s, 0] == n and fin[s] > 0
    b.close()
    b = Batch(gpu_ctx, geos, [0, 0, 0], leaf_size=2, grid=8)
    sums, fin, st = b.trace(m, precision=1)
    assert st["paths"] == 0 and not np.any(sums) and np.all(fin == -1.0)
    b.close()


@pytest.mark.gpu
def test_gpu_size_limits_and_invalid_arguments(gpu_ctx):
    geo = _film()
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=OM)
    cases = [
        (lambda: gpu_ctx.trace_grid(g, m, 300.0, 1, 10), "nangle"),
        (lambda: gpu_ctx.trace_grid(g, m, 0.0, 5, 10), "positive"),
        (lambda: gpu_ctx.trace_grid(g, m, 300.0, 5, -1), "np"),
        (lambda: gpu_ctx.trace_grid(g, m, 300.0, 5, 10, precision=7), "precision"),
        # 6 x 4 cells x 2^51 particles > 2^53 paths
        (lambda: gpu_ctx.trace_grid(g, m, 300.0, 5, 2 ** 51), r"2\^53 paths"),
        (lambda: gpu_ctx.replay(g, m, 300.0, 5, 10, 239, 2), "beyond the last cell"),
        (lambda: gpu_ctx.trace_log(g, m, 300.0, 5, 10, 0, 4, stride=0), "path range"),
        (lambda: gpu_ctx.trace_log(g, m, 300.0, 5, 10, 0, 4, max_points=0), "max_points"),
        (lambda: gpu_ctx.trace_alias(g, m, 10), "vel, mfp and weight"),
    ]
    for fn, msg in cases:
        with pytest.raises(_lib.PtransError, match=msg):
            fn()
    # oversized requests through the C-ABI with small buffers: refused before any write
    L = _lib.load()
    small = np.zeros(64, dtype=np.int64)
    st = _lib.PtStats()
    # 6 modes x (2^31 - 2) angles: more than 2^31 cells
    assert L.pt_trace_grid(gpu_ctx.handle, g.handle, m.handle, 300.0, 2 ** 31 - 1, None, 1, 0,
                           small.ctypes.data, ctypes.byref(st)) == -5
    assert "2^31 cells" in _lib.last_error()
    # 240 paths x 2^30 points x 24 B > 4 GiB
    assert L.pt_trace_log(gpu_ctx.handle, g.handle, m.handle, 300.0, 5, None, 10, 0, 1, 240, 0,
                          2 ** 30, small.ctypes.data, small.ctypes.data, None) == -1
    assert "4 GiB" in _lib.last_error()
    assert not np.any(small)
    # 2^30 surfaces: beyond the candidate-index range, refused before the arrays are read
    gdesc, _keep = geo.desc()
    gdesc.nsurface = 2 ** 30
    out = ctypes.c_void_p()
    assert L.pt_geom_create(gpu_ctx.handle, ctypes.byref(gdesc), ctypes.byref(out)) == -4
    assert "2^30" in _lib.last_error() and not out.value
    # the last path of the numbering is valid
    assert len(gpu_ctx.replay(g, m, 300.0, 5, 10, 239, 1)) == 1


@pytest.mark.gpu
def test_gpu_handles_of_another_context_are_refused(gpu_ctx):
    from ptrans.trace import Context
    geo = _film()
    with Context(seed=1, device=0) as other:
        g2 = other.geometry(geo)
        m = gpu_ctx.modes(omega=OM)
        with pytest.raises(_lib.PtransError, match="another context"):
            gpu_ctx.trace_grid(g2, m, 300.0, 5, 10)


def test_philox_fill_arguments():
    L = _lib.load()
    assert L.pt_philox_fill(1, 2, 0, -1, None) == -1
    assert L.pt_philox_fill(1, 2, 0, 0, None) == 0
    out = np.zeros(3)
    assert L.pt_philox_fill(7, 9, 4, 3, out.ctypes.data) == 0
    assert [float(x) for x in out] == [O.philox_draw(7, 9, d) for d in (4, 5, 6)]
