"""Likelihood-ratio adjoint of the alias-mode trace (an extension: the reference has no
derivatives).  d kappa / d theta = c E[(x - b) d log P(path) / d theta] with x = T/Lambda_m.

Pinned by central finite differences with common random numbers (same Philox streams,
Delta = 1e-2): the oracle's LR gradient against the oracle's FD on the CPU, the GPU's
against the oracle's batch sums, and the GPU's LR against the GPU's FD at C3 size
(12 inclusion specularities + 6 branch scattering-rate scales).
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptrans import gen, geometry, post
from ptrans.anphon import synthetic_table

M_TWOPI = 3.14159265358979323846 * 2.0
DELTA = 1.0e-2


def _modes(nb, nk, seed, mfp_range):
    t = synthetic_table(nb, nk, seed, mfp_min_nm=mfp_range[0], mfp_max_nm=mfp_range[1])
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    return omega, t.vel, t.mfp_nm, t.weight


def _weights(omega, vel, mfp, weight, T=300.0):
    cap = np.vectorize(lambda o: post.heat_capacity(float(o), T))(omega)
    return (cap * vel * mfp * weight).reshape(-1, order="F")


class OracleSetup:
    def __init__(self, model, nb, nk, seed_modes, mfp_range, pvals, group=None):
        self.geo = geometry.build_geometry(model)
        self.length = self.geo.box[0]
        self.gdesc, self.garr = self.geo.desc()
        if group is None:
            self.group, self.ngroup = geometry.inclusion_groups(self.geo)
        else:
            self.group, self.ngroup = np.asarray(group, dtype=np.int32), int(max(group) + 1)
        self.p0 = geometry.group_specularity(self.garr["surfaces"]["specularity"], self.group,
                                             pvals)
        self.garr["surfaces"]["specularity"] = self.p0
        self.omega, self.vel, self.mfp, self.weight = _modes(nb, nk, seed_modes, mfp_range)
        self.mdesc, self._m = geometry.modes_desc(self.omega, self.vel, self.mfp, self.weight)
        self.prob, self.alias = O.build_alias(_weights(self.omega, self.vel, self.mfp, self.weight))
        self.nb = nb

    def sums(self, seed, n, nbatch, baseline=0.0, scale=None, p=None, adjoint=True):
        if p is not None:
            self.garr["surfaces"]["specularity"] = p
        try:
            return O.adjoint_alias(self.gdesc, self.mdesc, self.mfp, self.prob, self.alias, seed,
                                   0, n, group=self.group if adjoint else None,
                                   ngroup=self.ngroup if adjoint else 0, baseline=baseline,
                                   nbatch=nbatch, scale=scale)
        finally:
            self.garr["surfaces"]["specularity"] = self.p0


def _fd(plus, minus, length):
    """Central difference of kappa/kappa_bulk per batch with common random numbers, and
    its batch-means standard error."""
    c = 0.75 * length
    n = plus[:, 0]
    assert np.array_equal(n, minus[:, 0])
    d = c * (plus[:, 1] - minus[:, 1]) / (2 * DELTA)
    r, se = post._ratio_of_sums(d, n)
    return float(r), float(se)


def _small_setup():
    """A 600 x 100 x 100 nm wire whose y walls and z walls form the two specularity groups
    (transmission depends strongly on both), two branches with mean free paths 100 nm -
    1 um."""
    model = gen.nanowire(length=600.0, p_wall=0.5, nphonon=1, nangle=3)
    nsurf = len(geometry.build_geometry(model).flatten()["surfaces"])
    group = [-1] * nsurf
    group[2] = group[3] = 0
    group[4] = group[5] = 1
    return OracleSetup(model, 2, 3, 11, (100.0, 1000.0), [0.35, 0.6], group=group)


def _inclusions_setup():
    model = gen.inclusions(seed=5, L=400.0, nsphere=2, ncyl=1, rs=(40.0, 80.0), rc=(30.0, 50.0),
                           kind="Gamma", value=1.0, p=0.5, nphonon=1, nangle=3)
    return OracleSetup(model, 2, 3, 11, (150.0, 1500.0), [0.3, 0.6, 0.45])


def test_oracle_adjoint_totals_match_alias_tallies():
    s = _small_setup()
    sums = s.sums(3, 4000, 16, adjoint=True)
    mp, mt = O.trace_alias(s.gdesc, s.mdesc, s.mfp, s.prob, s.alias, 3, 0, 4000)
    assert sums[:, 0].sum() == mp.sum()
    x = (mt / s.mfp).sum()
    assert abs(sums[:, 1].sum() - x) <= 1e-12 * abs(x)
    # baseline shifts the scored sums by -b * (sum of scores); with b = 0 vs b != 0 the
    # kappa columns are unchanged
    sb = s.sums(3, 4000, 16, baseline=0.002)
    np.testing.assert_array_equal(sb[:, :3], sums[:, :3])


def test_oracle_lr_gradient_matches_central_differences():
    """Both specularity groups and both branch scales: |LR - FD| within 3 combined
    standard errors (batch means, common random numbers)."""
    s = _small_setup()
    n, nb_ = 150000, 100
    pilot = s.sums(99, 5000, 1, adjoint=False)
    base = pilot[0, 1] / pilot[0, 0]
    g = post.alias_gradients(s.sums(7, n, nb_, baseline=base), s.length, s.ngroup)
    fails = []
    for k in range(s.ngroup + s.nb):
        if k < s.ngroup:
            vals = np.array([s.p0[s.group == j][0] for j in range(s.ngroup)])
            up, dn = vals.copy(), vals.copy()
            up[k] += DELTA
            dn[k] -= DELTA
            plus = s.sums(7, n, nb_, p=geometry.group_specularity(s.p0, s.group, up), adjoint=False)
            minus = s.sums(7, n, nb_, p=geometry.group_specularity(s.p0, s.group, dn), adjoint=False)
        else:
            b = k - s.ngroup
            up, dn = np.ones(s.nb), np.ones(s.nb)
            up[b] += DELTA
            dn[b] -= DELTA
            plus = s.sums(7, n, nb_, scale=up, adjoint=False)
            minus = s.sums(7, n, nb_, scale=dn, adjoint=False)
        fd, fse = _fd(plus, minus, s.length)
        tol = 3.0 * np.hypot(g.grad_stderr[k], fse)
        print("component", k, "LR", g.grad[k], "+-", g.grad_stderr[k], "FD", fd, "+-", fse)
        if abs(g.grad[k] - fd) > tol:
            fails.append((k, g.grad[k], fd, tol))
    assert not fails, fails


def test_oracle_flipped_score_sign_is_caught():
    """A planted mistake in the estimator (the score of the diffuse branch with the wrong
    sign, i.e. -score instead of score for every component) fails the FD comparison: the
    test above is sensitive to it."""
    s = _small_setup()
    n, nb_ = 60000, 60
    g = post.alias_gradients(s.sums(7, n, nb_), s.length, s.ngroup)
    k = int(np.argmax(np.abs(g.grad[:s.ngroup]) / g.grad_stderr[:s.ngroup]))
    vals = np.array([s.p0[s.group == j][0] for j in range(s.ngroup)])
    up, dn = vals.copy(), vals.copy()
    up[k] += DELTA
    dn[k] -= DELTA
    plus = s.sums(7, n, nb_, p=geometry.group_specularity(s.p0, s.group, up), adjoint=False)
    minus = s.sums(7, n, nb_, p=geometry.group_specularity(s.p0, s.group, dn), adjoint=False)
    fd, fse = _fd(plus, minus, s.length)
    assert abs(-g.grad[k] - fd) > 3.0 * np.hypot(g.grad_stderr[k], fse)


def test_adjoint_rejects_bad_arguments():
    import ctypes
    from ptrans import _lib
    L = _lib.load()
    assert L.pt_adjoint_alias(None, None, None, 10, 0, None, None, 0, 0.0, 8, None, None, None,
                              None) != 0
    buf = ctypes.create_string_buffer(256)
    L.pt_last_error(buf, 256)
    assert b"NULL" in buf.value


# ------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 2])
def test_gpu_adjoint_sums_equal_oracle(gpu_ctx, precision):
    """Per-path scores are the same arithmetic on both sides (bit-identical paths); the
    batch sums differ only by the order of the FP64 additions."""
    s = _small_setup()
    gpu_ctx.set_seed(7)
    g = gpu_ctx.geometry(s.geo)
    g.set_specularity(s.p0)
    m = gpu_ctx.modes(omega=s.omega, vel=s.vel, mfp=s.mfp, weight=s.weight)
    scale = np.array([0.8, 1.25])
    print("adjoint vs oracle, precision", precision, flush=True)
    mp, mt, gs, st = gpu_ctx.adjoint_alias(g, m, 20000, surface_group=s.group, ngroup=s.ngroup,
                                           baseline=0.001, nbatch=32, branch_scale=scale,
                                           precision=precision)
    os_ = s.sums(7, 20000, 32, baseline=0.001, scale=scale)
    np.testing.assert_array_equal(gs[:, 0], os_[:, 0])
    np.testing.assert_allclose(gs, os_, rtol=1e-11, atol=1e-14)
    assert st["paths"] == 20000


@pytest.mark.gpu
def test_gpu_lr_gradient_matches_fd_c3(gpu_ctx):
    """C3: 8 spheres + 4 z-cylinders in a 1000 nm cube, each inclusion with its own
    specularity; 12 specularity and 6 branch-scale derivatives of kappa/kappa_bulk by LR
    against central FD with common random numbers, FP32 trace."""
    model = gen.inclusions(seed=3, kind="Gamma", value=2.0, p=0.5, nphonon=1, nangle=3)
    geo = geometry.build_geometry(model)
    group, ngroup = geometry.inclusion_groups(geo)
    rng = np.random.default_rng(8)
    pvals = rng.uniform(0.1, 0.9, ngroup)
    p0 = geometry.group_specularity(geo.flatten()["surfaces"]["specularity"], group, pvals)
    omega, vel, mfp, weight = _modes(6, 40, 13, (100.0, 5000.0))
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    n, nbatch, prec = 2_000_000, 200, 1
    gpu_ctx.set_seed(1234)
    g.set_specularity(p0)
    _, _, pilot, _ = gpu_ctx.adjoint_alias(g, m, 200000, nbatch=1, precision=prec)
    base = pilot[0, 1] / pilot[0, 0]
    gpu_ctx.set_seed(4321)
    _, _, sums, st = gpu_ctx.adjoint_alias(g, m, n, surface_group=group, ngroup=ngroup,
                                           baseline=base, nbatch=nbatch, precision=prec)
    lr = post.alias_gradients(sums, geo.box[0], ngroup)
    print("adjoint %d paths in %.2f ms" % (n, st["kernel_ms"]), flush=True)

    def fwd(p=p0, scale=None):
        g.set_specularity(p)
        return gpu_ctx.adjoint_alias(g, m, n, nbatch=nbatch, branch_scale=scale, precision=prec)[2]

    within2, fails = 0, []
    for k in range(ngroup + 6):
        if k < ngroup:
            up, dn = pvals.copy(), pvals.copy()
            up[k] += DELTA
            dn[k] -= DELTA
            plus = fwd(p=geometry.group_specularity(p0, group, up))
            minus = fwd(p=geometry.group_specularity(p0, group, dn))
        else:
            up, dn = np.ones(6), np.ones(6)
            up[k - ngroup] += DELTA
            dn[k - ngroup] -= DELTA
            plus = fwd(scale=up)
            minus = fwd(scale=dn)
        fd, fse = _fd(plus, minus, geo.box[0])
        sig = np.hypot(lr.grad_stderr[k], fse)
        print("component %2d  LR % .5f +- %.5f   FD % .5f +- %.5f   %.2f sigma"
              % (k, lr.grad[k], lr.grad_stderr[k], fd, fse, abs(lr.grad[k] - fd) / sig), flush=True)
        within2 += abs(lr.grad[k] - fd) <= 2 * sig
        if abs(lr.grad[k] - fd) > 3 * sig:
            fails.append(k)
    g.set_specularity(p0)
    assert not fails, fails
    assert within2 >= 0.75 * (ngroup + 6)


def test_kappa_function_backward_uses_adjoint_gradient():
    """torch.autograd plumbing with a stand-in evaluator (host logic only)."""
    import torch
    from ptrans.autograd import KappaEval, KappaFunction

    def evaluate(p, s):
        return KappaEval(value=float((p ** 2).sum() + 3 * s.sum()), stderr=0.0, grad_p=2 * p,
                         grad_s=np.full(s.shape, 3.0), grad_p_stderr=np.zeros_like(p),
                         grad_s_stderr=np.zeros_like(s), paths=1, kernel_ms=0.0)

    p = torch.tensor([0.2, 0.7], dtype=torch.float64, requires_grad=True)
    s = torch.tensor([1.0, 1.5, 0.5], dtype=torch.float64, requires_grad=True)
    k = KappaFunction.apply(p, s, evaluate)
    (2.5 * k).backward()
    np.testing.assert_allclose(p.grad.numpy(), 2.5 * 2 * np.array([0.2, 0.7]))
    np.testing.assert_allclose(s.grad.numpy(), [7.5, 7.5, 7.5])
    assert abs(float(k.detach()) - (0.53 + 9.0)) < 1e-12


@pytest.mark.gpu
def test_gpu_kappa_model_and_screening(gpu_ctx):
    import torch
    from ptrans.autograd import KappaModel
    from ptrans import screening
    s_ = _small_setup()
    g = gpu_ctx.geometry(s_.geo)
    g.set_specularity(s_.p0)
    m = gpu_ctx.modes(omega=s_.omega, vel=s_.vel, mfp=s_.mfp, weight=s_.weight)
    km = KappaModel(gpu_ctx, g, m, s_.group, s_.ngroup, npaths=200000, nbatch=32, precision=0,
                    seed=5)
    p = torch.tensor([0.35, 0.6], dtype=torch.float64, requires_grad=True)
    s = torch.ones(2, dtype=torch.float64, requires_grad=True)
    k = km(p, s)
    k.backward()
    ev = km.last
    np.testing.assert_allclose(p.grad.numpy(), ev.grad_p)
    np.testing.assert_allclose(s.grad.numpy(), ev.grad_s)
    # the first evaluation (baseline 0) equals a direct adjoint call on the same stream
    gpu_ctx.set_seed(5)
    _, _, sums, _ = gpu_ctx.adjoint_alias(g, m, 200000, surface_group=s_.group, ngroup=s_.ngroup,
                                          nbatch=32, precision=0)
    direct = post.alias_gradients(sums, s_.length, s_.ngroup)
    assert abs(direct.ratio - float(k.detach())) <= 1e-12 * abs(direct.ratio)
    res = screening.screen(structures=2, steps=1, npaths=100000, nkpoint=20)
    assert len(res) == 2 and all(np.isfinite(r["kappa_ratio"]) for r in res)
    assert res[0]["kappa_ratio"] <= res[1]["kappa_ratio"]
