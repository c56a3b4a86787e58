"""The oracle (oracle/ptrans_oracle.c) against golden vectors written by the reference
itself (tests/golden/make_golden.py): per path the exit, the number of random draws,
their call sites, every trace.out position (REAL(4)); per cell the transmitted counts
of TransmissionFunction.  Both transcendental variants are pinned."""
import numpy as np
import pytest

from _golden import CASES, Case
from oracle import oracle as O

MATHS = [("own", O.MATH_OWN), ("libm", O.MATH_LIBM)]


@pytest.mark.parametrize("tag,math", MATHS)
@pytest.mark.parametrize("name", CASES)
def test_paths_match_reference(name, tag, math):
    c = Case(name)
    r = c.ref(tag)
    m = c.meta
    rec, traj, toff, sites, soff = O.replay_log(c.gdesc, c.mdesc, m["mfp"], m["nangle"],
                                                m["nphonon"], m["paths_seed"], 0, m["npaths"],
                                                math=math)
    np.testing.assert_array_equal((rec["end_side"] != 1).astype(np.int8), r["trans"])
    np.testing.assert_array_equal(rec["draws"], r["draws"])
    np.testing.assert_array_equal(rec["events"], np.diff(r["traj_off"]) - 1)
    np.testing.assert_array_equal(toff, r["traj_off"])
    np.testing.assert_array_equal(traj.astype(np.float32), r["traj"])
    np.testing.assert_array_equal(soff, r["site_off"])
    np.testing.assert_array_equal(sites, r["site"])
    assert (rec["flags"] & 2).sum() == 0            # no start-grain failures in these inputs


@pytest.mark.parametrize("tag,math", MATHS)
@pytest.mark.parametrize("name", CASES)
def test_grid_counts_match_reference(name, tag, math):
    c = Case(name)
    m = c.meta
    counts, stats = O.trace_grid(c.gdesc, c.mdesc, m["mfp"], m["nangle"], m["nphonon"],
                                 m["grid_seed"], math=math)
    np.testing.assert_array_equal(counts, c.ref(tag)["grid"])
    assert counts[:, :, -1].sum() == 0              # ia = nangle is never traced (RT:1576)
    assert stats[0] == m["nbranch"] * m["nkpoint"] * (m["nangle"] - 1) * m["nphonon"]


def test_planted_mistake_is_caught():
    """A wrong seed (standing in for any change of the random stream) must fail the
    comparison: the golden vectors are not satisfied by accident."""
    c = Case("c1_wire_gray")
    m = c.meta
    rec = O.replay(c.gdesc, c.mdesc, m["mfp"], m["nangle"], m["nphonon"], m["paths_seed"] + 1,
                   0, m["npaths"])
    assert not np.array_equal(rec["draws"], c.ref("own")["draws"])
