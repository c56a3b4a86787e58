"""kappa as a differentiable function of surface specularities and branch scattering-rate
scales, for torch.

    model = KappaModel(ctx, geom, modes, group, ngroup, npaths=2_000_000)
    p = torch.full((ngroup,), 0.5, requires_grad=True)
    s = torch.ones(nbranch, requires_grad=True)
    k = model(p, s)            # kappa / kappa_bulk (or W/mK with kappa_bulk given)
    k.backward()               # p.grad, s.grad from the likelihood-ratio adjoint

One forward call runs pt_adjoint_alias once: the same paths give kappa and its gradient
(there is no separate backward trace).  The reference computes kappa by enumeration and
has no derivatives (src/RTMCphon.f90:2645-2755); this is an extension.
"""
from __future__ import annotations

import dataclasses
from typing import Callable, Optional

import numpy as np
import torch

from . import geometry as _geometry
from . import post
from .constants import PREC_FP32_RETEST


@dataclasses.dataclass
class KappaEval:
    value: float
    stderr: float
    grad_p: np.ndarray
    grad_s: np.ndarray
    grad_p_stderr: np.ndarray
    grad_s_stderr: np.ndarray
    paths: int
    kernel_ms: float


class KappaFunction(torch.autograd.Function):
    """forward(p, s, evaluate) -> kappa; evaluate(p: ndarray, s: ndarray) -> KappaEval.
    backward returns grad_out * dkappa/dp and grad_out * dkappa/ds."""

    @staticmethod
    def forward(ctx, p, s, evaluate: Callable[[np.ndarray, np.ndarray], KappaEval]):
        ev = evaluate(p.detach().cpu().double().numpy(), s.detach().cpu().double().numpy())
        ctx.save_for_backward(torch.as_tensor(ev.grad_p, dtype=p.dtype, device=p.device),
                              torch.as_tensor(ev.grad_s, dtype=s.dtype, device=s.device))
        ctx.last_eval = ev
        return torch.as_tensor(ev.value, dtype=p.dtype, device=p.device)

    @staticmethod
    def backward(ctx, grad_out):
        gp, gs = ctx.saved_tensors
        return grad_out * gp, grad_out * gs, None


class KappaModel:
    """kappa(p, s) on one geometry and mode table through pt_adjoint_alias.

    group / ngroup: specularity group of each surface (-1: fixed), e.g. from
      geometry.inclusion_groups; p[g] is the specularity of every surface of group g.
    npaths, nbatch, precision: per evaluation.  crn=True reuses one Philox seed for every
      evaluation (common random numbers: smooth in p and s across calls).
    kappa_bulk: multiplies kappa/kappa_bulk (1.0 returns the ratio).
    The baseline of each evaluation is E[x] of the previous one (a constant for the
    call, so the gradient stays unbiased)."""

    def __init__(self, ctx, geom, modes, group, ngroup: int, npaths: int = 1_000_000,
                 nbatch: int = 64, precision: int = PREC_FP32_RETEST, seed: int = 1,
                 crn: bool = True, kappa_bulk: float = 1.0):
        self.ctx, self.geom, self.modes = ctx, geom, modes
        self.group = np.asarray(group, dtype=np.int32)
        self.ngroup = int(ngroup)
        self.npaths, self.nbatch, self.precision = int(npaths), int(nbatch), int(precision)
        self.seed, self.crn = int(seed), bool(crn)
        self.kappa_bulk = float(kappa_bulk)
        self.length = float(geom.geometry.box[0])
        self._p_fixed = geom.specularity
        self._baseline = 0.0
        self._calls = 0
        self.last: Optional[KappaEval] = None

    def evaluate(self, p: np.ndarray, s: np.ndarray) -> KappaEval:
        p = np.asarray(p, dtype=np.float64)
        s = np.asarray(s, dtype=np.float64)
        if p.shape != (self.ngroup,) or s.shape != (self.modes.nbranch,):
            raise ValueError("p needs %d entries and s %d" % (self.ngroup, self.modes.nbranch))
        self.geom.set_specularity(_geometry.group_specularity(self._p_fixed, self.group, p))
        self.ctx.set_seed(self.seed if self.crn else self.seed + self._calls)
        self._calls += 1
        _, _, sums, st = self.ctx.adjoint_alias(
            self.geom, self.modes, self.npaths, surface_group=self.group, ngroup=self.ngroup,
            baseline=self._baseline, nbatch=self.nbatch, branch_scale=s,
            precision=self.precision)
        g = post.alias_gradients(sums, self.length, self.ngroup)
        self._baseline = float(sums[:, 1].sum() / max(sums[:, 0].sum(), 1.0))
        kb = self.kappa_bulk
        self.last = KappaEval(value=kb * g.ratio, stderr=kb * g.stderr, grad_p=kb * g.grad_p,
                              grad_s=kb * g.grad_s, grad_p_stderr=kb * g.grad_stderr[:self.ngroup],
                              grad_s_stderr=kb * g.grad_stderr[self.ngroup:], paths=g.paths,
                              kernel_ms=st["kernel_ms"])
        return self.last

    def __call__(self, p: torch.Tensor, s: torch.Tensor) -> torch.Tensor:
        return KappaFunction.apply(p, s, self.evaluate)
