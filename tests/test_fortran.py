"""The Fortran binding (p-trans_amd/fortran/ptrans.f90, iso_c_binding).

CPU: the interoperable types have the C structs' sizes; the RTMC adapter
(ptrans_rtmc.f90, the drop-in for TransmissionFunction, src/RTMCphon.f90:1544-1604)
compiles against the reference's own module where that build exists.
GPU: a Fortran host (tests/fortran/ft_trace.f90) that reads a flattened geometry and
calls pt_trace_grid through the binding gets the same counts as the ctypes host.
"""
import os
import subprocess

import numpy as np
import pytest

from ptrans import gen, geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "p-trans_amd", "fortran")
BUILD = os.path.join(FDIR, "_build")


def _built(name):
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        pytest.skip("%s is not built (__graft_entry__.build())" % name)
    return path


def test_fortran_types_have_c_layout():
    from ptrans import _lib
    from ptrans.geometry import PtGeomDesc, PtModesDesc, SURFACE_DTYPE
    out = subprocess.run([_built("ft_sizes")], capture_output=True, text=True, timeout=60,
                         check=True).stdout.split()
    sizes = [int(x) for x in out]
    import ctypes
    assert sizes == [SURFACE_DTYPE.itemsize, ctypes.sizeof(PtGeomDesc),
                     ctypes.sizeof(PtModesDesc), ctypes.sizeof(_lib.PtStats),
                     _lib.RECORD_DTYPE.itemsize] == [184, 112, 56, 120, 80]


def test_rtmc_adapter_compiles_against_reference_module():
    mod = os.path.join(ROOT, "oracle", "_ref", "rtmc.mod")
    if not os.path.exists(mod):
        pytest.skip("no build of the reference module here (development container only)")
    r = subprocess.run(["make", "-B", "-C", FDIR, "adapter"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(os.path.join(BUILD, "ptrans_rtmc.o"))


def write_geometry_bin(path, geo):
    """The flattened geometry in ft_trace's stream layout (native byte order)."""
    arr = geo.flatten()
    ngt = len(arr["grain_offset"]) - 1
    hdr = np.array([geo.kind_grain, geo.ngrain, geo.igrain_fill, geo.nmaterial, ngt,
                    len(arr["surfaces"]), 0, 0], dtype=np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.int64(arr["nvertex"]).tobytes())
        f.write(np.array(list(geo.box) + list(geo.scope), dtype=np.float64).tobytes())
        f.write(arr["grain_offset"].astype(np.int32).tobytes())
        f.write(arr["grain_incident"].astype(np.int32).tobytes())
        f.write(arr["generatrix"].astype(np.float64).tobytes())
        f.write(arr["surfaces"].tobytes())
        if arr["nvertex"]:
            f.write(arr["vertices"][:3 * arr["nvertex"]].astype(np.float64).tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_fortran_host_counts_equal_ctypes_host(gpu_ctx, tmp_path, precision):
    exe = _built("ft_trace")
    model = gen.film_holes(seed=2, nphonon=2000, nangle=7, mfp=(300.0,))
    geo = geometry.build_geometry(model)
    om = np.linspace(1e12, 9e13, 6).reshape(2, 3)
    write_geometry_bin(tmp_path / "geom.bin", geo)
    with open(tmp_path / "modes.bin", "wb") as f:
        f.write(np.array([2, 3], dtype=np.int32).tobytes())
        f.write(np.float64(om.max()).tobytes())
        f.write(np.asfortranarray(om).reshape(-1, order="F").astype(np.float64).tobytes())
    seed = 91
    print("ft_trace precision", precision, flush=True)
    r = subprocess.run([exe, str(tmp_path / "geom.bin"), str(tmp_path / "modes.bin"),
                        str(tmp_path / "counts.bin"), str(seed), "300.0", "7", "2000",
                        str(precision)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ft_trace ok" in r.stdout, r.stdout + r.stderr
    fc = np.fromfile(tmp_path / "counts.bin", dtype=np.int64).reshape(2, 3, 7, order="F")
    gpu_ctx.set_seed(seed)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om, max_omega=float(om.max()))
    c, _ = gpu_ctx.trace_grid(g, m, 300.0, 7, 2000, precision=precision)
    np.testing.assert_array_equal(fc, c)
    assert fc.sum() > 0


@pytest.mark.gpu
def test_fortran_host_trace_out_equals_python_writer(gpu_ctx, tmp_path):
    """pt_trace_log through module ptrans, written by Fortran's own list-directed output
    (as the reference writes trace.out), equals post.trace_out_lines on the positions the
    ctypes host records: same paths, same text."""
    from ptrans import post
    exe = _built("ft_trace")
    model = gen.film_holes(seed=2, nphonon=200, nangle=7, mfp=(300.0,))
    geo = geometry.build_geometry(model)
    om = np.linspace(1e12, 9e13, 6).reshape(2, 3)
    write_geometry_bin(tmp_path / "geom.bin", geo)
    with open(tmp_path / "modes.bin", "wb") as f:
        f.write(np.array([2, 3], dtype=np.int32).tobytes())
        f.write(np.float64(om.max()).tobytes())
        f.write(np.asfortranarray(om).reshape(-1, order="F").astype(np.float64).tobytes())
    seed = 17
    r = subprocess.run([exe, str(tmp_path / "geom.bin"), str(tmp_path / "modes.bin"),
                        str(tmp_path / "counts.bin"), str(seed), "300.0", "7", "200", "0",
                        str(tmp_path / "trace.out"), "30", "7"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ft_trace ok" in r.stdout, r.stdout + r.stderr
    gpu_ctx.set_seed(seed)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=om, max_omega=float(om.max()))
    tr, _ = gpu_ctx.trace_log(g, m, 300.0, 7, 200, 0, 30, stride=7)
    with open(tmp_path / "trace.out") as f:
        text = f.read()
    assert text == "\n".join(post.trace_out_lines(tr)) + "\n"
