"""Bulk phonon table: the .anphon format read by loadAnphon (src/RTMCphon.f90:661-733),
a reader with the same semantics, and the seeded synthetic tables used as inputs.

Layout: arrays are (nbranch, nkpoint) like Anphon_type (RT:161-174); omega is stored in
rad/s after the reader's conversion omega[cm^-1] * 0.03 * 2*pi * 1e12 (RT:727).
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

from .constants import HBAR, KB, M_TWOPI


@dataclasses.dataclass
class Anphon:
    material: str
    rtemp: float
    vol: float
    omega: np.ndarray      # (nb, nk) rad/s
    tau: np.ndarray        # (nb, nk)
    vel: np.ndarray        # (nb, nk)
    mfp: np.ndarray        # (nb, nk)
    weight: np.ndarray     # (nb, nk)

    @property
    def nbranch(self) -> int:
        return self.omega.shape[0]

    @property
    def nkpoint(self) -> int:
        return self.omega.shape[1]

    @property
    def max_omega(self) -> float:
        """anphon%maxOmega = maxval(omega) (RT:728)."""
        return float(np.max(self.omega))


def _words_then(line: str, nskip: int, conv):
    toks = line.replace(",", " ").split()
    if len(toks) <= nskip:
        raise ValueError(f"broken .anphon header line: {line!r}")
    return conv(toks[nskip])


def read_anphon(path: str, material: str | None = None) -> Anphon:
    """loadAnphon (RT:661-733): two free lines, RTemp after 5 words, vol after 3 words,
    nkpoint and nbranch after 4 words, one free line, then nk x nb rows
    "word word omega tau vel mfp weight" with the branch index fastest."""
    with open(path) as f:
        lines = f.read().splitlines()
    rtemp = _words_then(lines[2], 5, float)
    vol = _words_then(lines[3], 3, float)
    nk = _words_then(lines[4], 4, int)
    nb = _words_then(lines[5], 4, int)
    rows = []
    for ln in lines[7:]:
        if ln.strip():
            rows.append(ln.split())
        if len(rows) == nk * nb:
            break
    if len(rows) != nk * nb:
        raise ValueError("broken .anphon table: expected %d rows, got %d" % (nk * nb, len(rows)))
    data = np.array([[float(x) for x in r[2:7]] for r in rows], dtype=np.float64)
    data = data.reshape(nk, nb, 5)            # row order: ik outer, ib inner
    omega_cm = np.ascontiguousarray(data[:, :, 0].T)
    omega = omega_cm * 0.03 * M_TWOPI * 1.0e12   # RT:727, evaluated left to right
    return Anphon(material=material or os.path.splitext(os.path.basename(path))[0],
                  rtemp=rtemp, vol=vol, omega=omega,
                  tau=np.ascontiguousarray(data[:, :, 1].T),
                  vel=np.ascontiguousarray(data[:, :, 2].T),
                  mfp=np.ascontiguousarray(data[:, :, 3].T),
                  weight=np.ascontiguousarray(data[:, :, 4].T))


def heat_capacity(omega: np.ndarray, temperature: float) -> np.ndarray:
    """Mode heat capacity k_B ((x/2)/sinh(x/2))^2, x = hbar omega/(k_B T) (RT:2696-2701)."""
    x = HBAR * omega / (KB * temperature)
    h = x / 2.0
    out = np.empty_like(omega)
    small = h == 0.0
    out[small] = KB
    out[~small] = KB * (h[~small] / np.sinh(h[~small])) ** 2
    return out


@dataclasses.dataclass
class SyntheticTable:
    """The seeded synthetic bulk table: f_k = (k-1/2)*fmax/K THz, v ~ U(1,9) km/s,
    mfp ~ logU(1 nm, 10 um), tau = mfp/v, weight 1.  Written so that loadAnphon reads it
    back unchanged: omega in cm^-1 = f[THz]/0.03, tau in ps, v in m/s, mfp in nm."""
    nbranch: int
    nkpoint: int
    f_thz: np.ndarray      # (nb, nk)
    vel: np.ndarray        # (nb, nk) m/s
    mfp_nm: np.ndarray     # (nb, nk)
    tau_ps: np.ndarray     # (nb, nk)
    weight: np.ndarray     # (nb, nk)
    rtemp: float = 300.0
    vol: float = 1.0e-27

    def omega_cm(self) -> np.ndarray:
        return self.f_thz / 0.03


def synthetic_table(nbranch: int, nkpoint: int, seed: int, fmax_thz: float = 15.0,
                    vmin: float = 1.0e3, vmax: float = 9.0e3,
                    mfp_min_nm: float = 1.0, mfp_max_nm: float = 1.0e4) -> SyntheticTable:
    rng = np.random.default_rng(seed)
    k = np.arange(1, nkpoint + 1, dtype=np.float64)
    f = np.broadcast_to((k - 0.5) * fmax_thz / nkpoint, (nbranch, nkpoint)).copy()
    vel = rng.uniform(vmin, vmax, size=(nbranch, nkpoint))
    mfp = np.exp(rng.uniform(np.log(mfp_min_nm), np.log(mfp_max_nm), size=(nbranch, nkpoint)))
    tau = mfp * 1e-9 / vel * 1e12
    return SyntheticTable(nbranch, nkpoint, f, vel, mfp, tau, np.ones((nbranch, nkpoint)))


def write_anphon(path: str, t: SyntheticTable) -> None:
    """Write a table in the layout loadAnphon reads (RT:679-726).  Values are written with
    17 significant digits so the reader recovers the same doubles."""
    with open(path, "w") as f:
        f.write("# synthetic bulk phonon table for the ptrans trace\n")
        f.write("# columns: k b omega[cm^-1] tau[ps] v[m/s] mfp[nm] multiplicity\n")
        f.write("reference temperature in units kelvin %.17g\n" % t.rtemp)
        f.write("cell volume m3 %.17g\n" % t.vol)
        f.write("number of k points %d\n" % t.nkpoint)
        f.write("number of phonon branches %d\n" % t.nbranch)
        f.write("#\n")
        om = t.omega_cm()
        for ik in range(t.nkpoint):
            for ib in range(t.nbranch):
                f.write("%d %d %.17g %.17g %.17g %.17g %.17g\n" % (
                    ik + 1, ib + 1, om[ib, ik], t.tau_ps[ib, ik], t.vel[ib, ik],
                    t.mfp_nm[ib, ik], t.weight[ib, ik]))
