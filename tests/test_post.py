"""effectiveMFP and fileoutput (src/RTMCphon.f90:2418-2451, 2601-2881) on the host,
against the reference's own routines run on the same transmission tables."""
import os

import numpy as np
import pytest

from _golden import GOLDEN
from ptrans import anphon, mcp, post

POST = ["post_matthiessen", "post_mfp_log"]


def _load(name):
    d = os.path.join(GOLDEN, name)
    model = mcp.read_mcp(os.path.join(d, "input.mcp"))
    table = anphon.read_anphon(os.path.join(d, "Si.anphon"), "Si")
    counts = np.load(os.path.join(d, "counts.npy"))
    return d, model, table, counts


@pytest.mark.parametrize("name", POST)
def test_effective_mfp_bitexact(name):
    d, model, table, counts = _load(name)
    ang = post.vangle(model.nangle)
    ref = np.load(os.path.join(d, "semfp_ref.npy"))
    for s in range(counts.shape[-1]):
        t = counts[..., s].astype(np.float64) / float(model.nphonon)
        np.testing.assert_array_equal(post.effective_mfp(t, ang, model.box[0]), ref[:, :, s])


@pytest.mark.parametrize("name", POST)
def test_output_files_match_reference(name, tmp_path):
    d, model, table, counts = _load(name)
    ang = post.vangle(model.nangle)
    semfp = np.stack([post.effective_mfp(counts[..., s] / float(model.nphonon), ang, model.box[0])
                      for s in range(counts.shape[-1])], axis=-1)
    res = post.kappa(semfp, table, model.temperature, model.mean_free_paths(),
                     matthiessen=model.mfp is None)
    post.write_outputs(res, str(tmp_path), model.nphonon, semfp, model.temperature,
                       table.nbranch, table.nkpoint)
    for fn in ("thermalproperty.out", "effectiveMFP.out", "phononconductivity.out"):
        with open(os.path.join(d, fn)) as f:
            ref = [ln.rstrip() for ln in f.read().splitlines()]
        with open(os.path.join(str(tmp_path), fn)) as f:
            got = [ln.rstrip() for ln in f.read().splitlines()]
        assert got == ref, fn


def test_kappa_linear_in_transmission_and_stderr():
    """kappa is linear in T in the limit of Matthiessen with B << Lambda: scaling T by
    1/2 halves SEMFP exactly; the binomial standard error vanishes for T in {0, 1}."""
    d, model, table, counts = _load("post_matthiessen")
    ang = post.vangle(model.nangle)
    t = counts[..., 0] / float(model.nphonon)
    a = post.effective_mfp(t, ang, model.box[0])
    b = post.effective_mfp(t * 0.5, ang, model.box[0])
    np.testing.assert_allclose(b, 0.5 * a, rtol=1e-15)
    full = np.full_like(counts[..., 0], model.nphonon)
    full[:, :, -1] = 0
    semfp = post.effective_mfp(full / float(model.nphonon), ang, model.box[0])[..., None]
    res = post.kappa(semfp, table, model.temperature, [1e10], True, counts=full,
                     np_=model.nphonon, angles=ang, boxx=model.box[0])
    assert res.kappa_stderr == 0.0
    # T = 1 for every angle: Simpson of sin*cos over [0, pi/2] is ~1/2, so SEMFP ~ 3L/4
    np.testing.assert_allclose(semfp, 0.75 * model.box[0], rtol=2e-3)


def test_fortran_e_format():
    assert post.fortran_e(12.5, 15, 6) == "   0.125000E+02"
    assert post.fortran_e(-0.00123, 15, 6) == "  -0.123000E-02"
    assert post.fortran_e(0.0, 15, 6) == "   0.000000E+00"
    assert post.fortran_e(9.9999999e5, 15, 6) == "   0.100000E+07"
