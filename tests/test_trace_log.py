"""Trajectory recorder: pt_trace_log and the trace.out writer (RT:1661-1665, 1818-1831,
3377-3382) against the reference's own trace.out text (tests/golden/<case>/trace_ref.out,
the reference linked with the project's transcendentals, first 40 paths of the case)."""
import os

import numpy as np
import pytest

from _golden import Case
from oracle import oracle as O
from ptrans import post

TEXT_CASES = ["c1_wire_gray", "c4_voronoi_gamma"]


def _ref_text(case):
    with open(os.path.join(case.dir, "trace_ref.out")) as f:
        return f.read()


def _blocks(text):
    out = []
    for part in text.split("\n\n\n"):
        rows = [[float(t) for t in ln.split()] for ln in part.split("\n") if ln.strip()]
        if rows:
            out.append(np.array(rows, dtype=np.float32))
    return out


@pytest.mark.parametrize("name", TEXT_CASES)
def test_writer_reproduces_reference_text(name):
    """The reference's REAL(4) values, written back by post.trace_out_lines, give the
    reference's file byte for byte (flang list-directed formatting)."""
    text = _ref_text(Case(name))
    blocks = _blocks(text)
    assert len(blocks) == 40
    assert "\n".join(post.trace_out_lines(blocks)) + "\n" == text


@pytest.mark.parametrize("name", TEXT_CASES)
def test_oracle_trajectories_write_the_reference_file(name):
    c = Case(name)
    m = c.meta
    _, traj, toff, _, _ = O.replay_log(c.gdesc, c.mdesc, m["mfp"], m["nangle"], m["nphonon"],
                                       m["paths_seed"], 0, 40, math=O.MATH_OWN)
    blocks = [traj[toff[i]:toff[i + 1]] for i in range(40)]
    assert "\n".join(post.trace_out_lines(blocks)) + "\n" == _ref_text(c)


def test_writer_header_and_fp64_option():
    lines = post.trace_out_lines([np.array([[0.0, 1.5, 2.0], [3.0, 4.0, 5.0]])], 1000, 25)
    assert lines[:4] == ["# Datatype = trace", "# particle_trace =     1000",
                         "# interval_trace =       25", ""]
    assert lines[4:] == [" 0. 1.5 2.", " 3. 4. 5.", "", ""]
    l64 = post.trace_out_lines([np.array([[0.1, 1.0 / 3.0, 2.0]])], fp64=True)
    assert [float(x) for x in l64[0].split()] == [0.1, 1.0 / 3.0, 2.0]


@pytest.mark.gpu
@pytest.mark.parametrize("name", TEXT_CASES)
def test_gpu_trace_log_equals_oracle_and_reference_text(gpu_ctx, name):
    c = Case(name)
    m = c.meta
    gpu_ctx.set_seed(m["paths_seed"])
    g = gpu_ctx.geometry(c.geom)
    t = c.table
    mm = gpu_ctx.modes(omega=t.omega)
    print("case", name, flush=True)
    rec0, traj, toff, _, _ = O.replay_log(c.gdesc, c.mdesc, m["mfp"], m["nangle"], m["nphonon"],
                                          m["paths_seed"], 0, 40, math=O.MATH_OWN)
    for prec in (0, 2):
        tr, rec = gpu_ctx.trace_log(g, mm, m["mfp"], m["nangle"], m["nphonon"], 0, 40,
                                    max_points=4096, precision=prec)
        for i in range(40):
            np.testing.assert_array_equal(tr[i].view(np.uint64),
                                          traj[toff[i]:toff[i + 1]].view(np.uint64))
        np.testing.assert_array_equal(rec["events"], rec0["events"])
        assert "\n".join(post.trace_out_lines(tr)) + "\n" == _ref_text(c)
    # stride and truncation: every 3rd path, at most 4 points each
    tr3, _ = gpu_ctx.trace_log(g, mm, m["mfp"], m["nangle"], m["nphonon"], 0, 13, stride=3,
                               max_points=4)
    for i in range(13):
        full = traj[toff[3 * i]:toff[3 * i + 1]]
        np.testing.assert_array_equal(tr3[i], full[:4])
