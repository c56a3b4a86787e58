"""Surface-law set-up of the .mcp input: GammaDeterm (TBC -> gamma) and the
FluctuationSpecularity jitter.

gamma_determ(table, tbc, temperature) restates GammaDeterm (src/RTMCphon.f90:2499-2599,
ModeNum = 1): Newton's method on gamma for the thermal boundary conductance of the
Gamma law t = 1 / (gamma omega / omega_max + 1),

    G(gamma) = S(gamma) / (1 - S(gamma) / S_a),
    S(gamma) = sum_{k, b} w C v t / 2 / vol / Nq,   S_a = S with t = 1,
    C = k_B ((x/2) / sinh(x/2))^2,  x = hbar omega / (k_B T)   (C = k_B for x <= 1e-10),
    Nq = int(sum w / nb),

starting at gamma = 1e-2, at most 10000 steps, stopping when |G - TBC| / G' <= 1e-6, with
the sums taken in the reference's order (k outer, b inner) so the result is bit-for-bit
the reference's.  The .mcp gives TBC in MW/m^2/K; RT:866 multiplies by 1e6.

fluctuate_specularity(p, sigma, seed, first_draw) is normal_distribution (RT:1190-1201)
applied at RT:1151-1157: p + (sum of 12 uniforms - 6) sigma, clipped to [0, 1], with the
uniforms of the set-up Philox stream (PT_SETUP_STREAM) instead of the runtime's
RANDOM_NUMBER.  Surfaces draw in the reference's order: the internal facets of grains
1..ngrain in file order (fill-grain copies and box walls are not jittered, RT:1389).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from .constants import HBAR, KB   # RT:67-68

SETUP_STREAM = 1 << 62       # PT_SETUP_STREAM in ptrans.h


class GammaError(ValueError):
    pass


def _cap(omega: float, temperature: float) -> float:
    x = HBAR * omega / (KB * temperature)
    if x <= 1.0e-10:
        return KB
    return KB * ((x / 2) / math.sinh(x / 2)) ** 2


def gamma_determ(table, tbc: float, temperature: float, max_iter: int = 10000,
                 eps: float = 1.0e-6) -> float:
    """gamma of the Gamma law whose boundary conductance is tbc (W/m^2/K) for the bulk
    table (an anphon.Anphon: omega (nb, nk) rad/s, vel, weight, vol) at temperature (K).
    Raises GammaError where the reference reports "Gamma cannot determine" (gamma < 0)."""
    om = np.asarray(table.omega, dtype=np.float64)
    nb, nk = om.shape
    vel = np.asarray(table.vel, dtype=np.float64)
    w = np.asarray(table.weight, dtype=np.float64)
    vol = float(table.vol)
    omax = float(om.max())
    nqfbz = int(float(np.sum(w)) / nb)
    # element order of the reference's loops: jj = 1..nk outer, kk = 1..nb inner
    order = [(kk, jj) for jj in range(nk) for kk in range(nb)]
    cap = {(kk, jj): _cap(float(om[kk, jj]), temperature) for kk, jj in order}
    g = 1.0e-2
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        g = _newton(order, om, omax, w, cap, vel, vol, nqfbz, tbc, g, max_iter, eps)
    g = float(g)
    if g < 0:
        raise GammaError("Gamma cannot determine (TBC %g W/m^2/K gives gamma %g)" % (tbc, g))
    return g


def _newton(order, om, omax, w, cap, vel, vol, nqfbz, tbc, g, max_iter, eps):
    for _ in range(max_iter):
        s = sa = ds = 0.0
        for kk, jj in order:
            h = 1 / (g * (om[kk, jj] / omax) + 1)
            s = s + w[kk, jj] * cap[kk, jj] * vel[kk, jj] * h / 2 / vol / nqfbz
        for kk, jj in order:
            sa = sa + w[kk, jj] * cap[kk, jj] * vel[kk, jj] * 1 / 2 / vol / nqfbz
        for kk, jj in order:
            h = -(om[kk, jj] / omax) / (g * (om[kk, jj] / omax) + 1) ** 2
            ds = ds + w[kk, jj] * cap[kk, jj] * vel[kk, jj] * h / 2 / vol / nqfbz
        f = s / (1 - s / sa) - tbc
        df = ds / (1 - s / sa) ** 2
        step = f / df
        if abs(step) <= eps:
            break
        g = g - step
    return g


def boundary_conductance(table, gamma: float, temperature: float) -> float:
    """G(gamma) of the Gamma law (W/m^2/K), the quantity gamma_determ inverts."""
    om = np.asarray(table.omega, dtype=np.float64)
    nb = om.shape[0]
    w = np.asarray(table.weight, dtype=np.float64)
    c = np.vectorize(lambda o: _cap(o, temperature))(om)
    base = w * c * np.asarray(table.vel, dtype=np.float64) / 2 / float(table.vol) / int(w.sum() / nb)
    s = float(np.sum(base / (gamma * om / om.max() + 1)))
    sa = float(np.sum(base))
    return s / (1 - s / sa)


def setup_uniforms(seed: int, first_draw: int, n: int) -> np.ndarray:
    """Draws first_draw .. first_draw + n - 1 of the set-up stream (through libptrans)."""
    from . import _lib
    out = np.zeros(n, dtype=np.float64)
    _lib.check(_lib.load().pt_philox_fill(ctypes.c_uint64(seed), ctypes.c_uint64(SETUP_STREAM),
                                          ctypes.c_uint64(first_draw), n, out.ctypes.data))
    return out


def fluctuate_specularity(p: float, sigma: float, u12) -> float:
    """normal_distribution(0, sigma) added to p and clipped (RT:1151-1157, 1190-1201)."""
    acc = 0.0
    for u in u12:
        acc = acc + float(u)
    p = p + ((acc - 6.0) * sigma + 0.0)
    if p < 0.0:
        p = 0.0
    if p > 1.0:
        p = 1.0
    return p
