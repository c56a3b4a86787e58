"""From transmission to effective mean free paths and thermal conductivity: the host
steps after the trace, effectiveMFP (src/RTMCphon.f90:2418-2451) and fileoutput
(RT:2601-2881), with the reference's arithmetic order and output formats, plus the
standard error of kappa that the reference does not report.

Units as in the reference: SEMFP in nm, bulk and effective MFP inside fileoutput in m,
kappa in W m^-1 K^-1 when the table's volume is in m^3.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional

import numpy as np

from .constants import HBAR, KB, M_PI_2


def vangle(nangle: int) -> np.ndarray:
    """param%vangle, RT:1554-1556."""
    return np.array([0.0 + float(ia - 1) / float(nangle - 1) * M_PI_2 for ia in range(1, nangle + 1)])


def effective_mfp(transmit: np.ndarray, angles: np.ndarray, boxx: float) -> np.ndarray:
    """effectiveMFP (RT:2418-2451): composite Simpson of sin(t)cos(t)T(t) over the
    angle grid, times (3/2)L.  transmit is (nb, nk, nangle); returns EMFP (nk, nb) in the
    reference's layout (SEMFP(:,:,s)).  The shared WT array of the reference's OpenMP
    loop (RT:2423, 2427) is a data race; this is its single-threaded meaning."""
    nb, nk, na = transmit.shape
    ds = (angles[1] - angles[0]) / 3.0 * 3.0 * 0.5 * boxx
    sc = np.array([math.sin(a) * math.cos(a) for a in angles])
    out = np.zeros((nk, nb))
    for ik in range(nk):
        for ib in range(nb):
            wt = sc * transmit[ib, ik, :]
            acc = 0.0
            for ia in range(1, na - 1, 2):            # ia = 2, 4, ..., nangle-1 (1-based)
                acc = acc + (wt[ia - 1] + 4.0 * wt[ia] + wt[ia + 1])
            out[ik, ib] = acc * ds
    return out


def heat_capacity(omega: float, temperature: float) -> float:
    """cap(ii,jj), RT:2696-2701."""
    x = HBAR * omega / (KB * temperature)
    if x <= 1.0e-10:
        return KB
    return KB * ((x / 2) / math.sinh(x / 2)) ** 2


@dataclasses.dataclass
class KappaResult:
    kappa: float                 # sum of Thermal_conductivity (RT:2754-2755)
    mode_kappa: np.ndarray       # (nb, nk) after the reference's per-branch sort by omega
    emfp_m: np.ndarray           # (nb, nk) effective MFP [m]
    bulk_mfp_m: np.ndarray       # (nb, nk)
    omega: np.ndarray            # (nb, nk) sorted
    kappa_stderr: Optional[float] = None


def kappa(semfp: np.ndarray, table, temperature: float, mfp_samples: List[float],
          matthiessen: bool, counts: Optional[np.ndarray] = None, np_: int = 0,
          angles: Optional[np.ndarray] = None, boxx: float = 0.0) -> KappaResult:
    """fileoutput's conductivity (RT:2645-2755) for nmaterial = 1.

    semfp: (nk, nb, nMFP) in nm.  table: an Anphon (omega rad/s, tau ps, vel, weight).
    If counts (nb, nk, nangle[, nMFP]) and np_ are given, also the standard error of
    kappa from the binomial variance T(1-T)/np of every traced cell, propagated
    linearly through Simpson and the kappa sum (Matthiessen: to first order)."""
    nb, nk = table.omega.shape
    rtemp = table.rtemp
    tau = table.tau.copy()
    vel = table.vel.copy()
    omega = table.omega.copy()
    weight = table.weight.copy()
    semfp = semfp.copy()
    mfp = np.zeros((nb, nk))
    for ii in range(nk):
        for jj in range(nb):
            tau[jj, ii] = tau[jj, ii] * 1.0e-12 * (rtemp / temperature)
            mfp[jj, ii] = tau[jj, ii] * vel[jj, ii]
    nqfbz = int(weight.sum(dtype=np.float64) / nb)
    # per-branch exchange sort by omega (RT:2658-2691)
    for ii in range(nb):
        if np.all(omega[ii, :-1] <= omega[ii, 1:]):
            continue                    # already ordered: the exchange sort swaps nothing
        for jj in range(nk - 1):
            for kk in range(jj + 1, nk):
                if omega[ii, jj] > omega[ii, kk]:
                    for a in (omega, weight, tau, vel, mfp):
                        a[ii, jj], a[ii, kk] = a[ii, kk], a[ii, jj]
                    semfp[jj, ii, :], semfp[kk, ii, :] = semfp[kk, ii, :].copy(), semfp[jj, ii, :].copy()
    cap = np.array([[heat_capacity(omega[ii, jj], temperature) for jj in range(nk)] for ii in range(nb)])
    emfp = np.zeros((nb, nk))
    tc = np.zeros((nb, nk))
    vm = list(mfp_samples)
    for ii in range(nb):
        for jj in range(nk):
            if matthiessen:
                if mfp[ii, jj] == 0.0:
                    emfp[ii, jj] = semfp[jj, ii, 0] * 1.0e-9
                else:
                    b = semfp[jj, ii, 0] * 1.0e-9
                    inv_b = 1.0 / b if b != 0.0 else math.inf
                    emfp[ii, jj] = 1.0 / (1.0 / mfp[ii, jj] + inv_b)
            else:
                m = mfp[ii, jj]
                if m < vm[0] * 1.0e-9:
                    emfp[ii, jj] = semfp[jj, ii, 0] * 1.0e-9 * m / (vm[0] * 1.0e-9)
                elif m > vm[-1] * 1.0e-9:
                    emfp[ii, jj] = semfp[jj, ii, len(vm) - 1] * 1.0e-9
                else:
                    for s in range(len(vm) - 1):
                        if vm[s] * 1.0e-9 <= m < vm[s + 1] * 1.0e-9:
                            t = (m - vm[s] * 1.0e-9) / ((vm[s + 1] - vm[s]) * 1.0e-9)
                            emfp[ii, jj] = (semfp[jj, ii, s] * 1.0e-9 * (1.0 - t)
                                            + semfp[jj, ii, s + 1] * 1.0e-9 * t)
                            break
            tc[ii, jj] = cap[ii, jj] * weight[ii, jj] * emfp[ii, jj] * vel[ii, jj] / 3 / table.vol / nqfbz
    total = 0.0
    for jj in range(nk):                 # column-major order of sum(Thermal_conductivity)
        for ii in range(nb):
            total = total + tc[ii, jj]
    res = KappaResult(kappa=total, mode_kappa=tc, emfp_m=emfp, bulk_mfp_m=mfp, omega=omega)
    if counts is not None and np_ > 0 and angles is not None and matthiessen:
        res.kappa_stderr = _kappa_stderr_matthiessen(counts, np_, angles, boxx, semfp, mfp,
                                                     cap, weight, vel, table.vol, nqfbz,
                                                     table.omega, omega)
    return res


def _kappa_stderr_matthiessen(counts, np_, angles, boxx, semfp, mfp, cap, weight, vel, vol,
                              nqfbz, omega_unsorted, omega_sorted):
    """sigma_kappa with independent binomial cells: Var T = T(1-T)/np; SEMFP is linear in
    T (Simpson weights w_a), and d kappa_m / d SEMFP_m = c_m * (EMFP_m/B_m)^2 * 1e-9
    with B = SEMFP*1e-9 (Matthiessen)."""
    nb, nk, na = counts.shape
    T = counts / float(np_)
    var_t = T * (1.0 - T) / float(np_)
    ds = (angles[1] - angles[0]) / 3.0 * 3.0 * 0.5 * boxx
    sc = np.sin(angles) * np.cos(angles)
    w = np.zeros(na)
    for ia in range(1, na - 1, 2):
        w[ia - 1] += 1.0
        w[ia] += 4.0
        w[ia + 1] += 1.0
    w = w * sc * ds
    var_semfp = (var_t * w ** 2).sum(axis=2)          # (nb, nk), unsorted order
    total = 0.0
    for ii in range(nb):
        order = np.argsort(omega_unsorted[ii], kind="stable")
        for jj in range(nk):
            b = semfp[jj, ii, 0] * 1.0e-9
            m = mfp[ii, jj]
            if b <= 0.0 or m == 0.0:
                continue
            e = 1.0 / (1.0 / m + 1.0 / b)
            c = cap[ii, jj] * weight[ii, jj] * vel[ii, jj] / 3 / vol / nqfbz
            d = c * (e / b) ** 2 * 1.0e-9
            total += d * d * var_semfp[ii, order[jj]]
    return math.sqrt(total)


# ------------------------------------------------------------------ output files

def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d: 0.ddddE+xx, right-justified in w columns."""
    if x == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        t = "%.*e" % (d - 1, abs(x))
        mant, ex = t.split("e")
        digits = mant.replace(".", "")
        e = int(ex) + 1
        s = ("-" if x < 0 else "") + "0." + digits + ("E%+03d" % e)
    return s.rjust(w)


def write_outputs(res: KappaResult, directory: str, nphonon: int, semfp: np.ndarray,
                  temperature: float, nbranch: int, nkpoint: int) -> None:
    """thermalproperty.out, effectiveMFP.out, phononconductivity.out (RT:2634-2871)."""
    ave = float(semfp.sum()) / (nbranch * nkpoint * semfp.shape[2])
    with open(os.path.join(directory, "thermalproperty.out"), "w") as f:
        f.write("# Datatype = thermalproperty\n")
        f.write("# Incident_Phonon                               =   %6d\n" % nphonon)
        f.write("# Averaged Effective_MeanFreePath [nm]          = %20.10f\n" % ave)
        f.write("# Thermal_conductivity@%sK [W m-1 K-1]  = %20.10f\n"
                % (("%.6f" % temperature), res.kappa))
    nb, nk = res.mode_kappa.shape
    recs = [(ii + 1, jj + 1, res.bulk_mfp_m[ii, jj], res.emfp_m[ii, jj], res.mode_kappa[ii, jj])
            for ii in range(nb) for jj in range(nk)]
    n = len(recs)
    for i in range(n):                                      # RT:2775-2784
        for j in range(i + 1, n):
            if recs[i][3] > recs[j][3]:
                recs[i], recs[j] = recs[j], recs[i]
    acc = []
    for i, r in enumerate(recs):
        acc.append(r[4] if i == 0 else r[4] + acc[-1])
    with open(os.path.join(directory, "effectiveMFP.out"), "w") as f:
        f.write("# Datatype = Effective Mean Free Path of phonon.\n")
        f.write("# data are sorted by effective MFP in ascending order.\n")
        f.write("#   total branches: %3d\n" % nb)
        f.write("#   total Kpoints:  %3d\n" % nk)
        f.write("#\n")
        f.write("# branch, kpoint, bulk MFP[nm], effective MFP[nm], ThermalConductivity, "
                "and its accumulated[W m-1 K-1].\n")
        for r, a in zip(recs, acc):
            f.write("%5d%5d%s%s%s%s\n" % (r[0], r[1], fortran_e(r[2] * 1.0e9, 15, 6),
                                          fortran_e(r[3] * 1.0e9, 15, 6), fortran_e(r[4], 15, 6),
                                          fortran_e(a, 15, 6)))
    freqs = []
    for ii in range(nb):
        for jj in range(nk):
            o = res.omega[ii, jj]
            if o not in freqs:
                freqs.append(o)
    freqs.sort()
    nf = len(freqs) - 1                                     # the last one is dropped (RT:2841)
    spec = np.zeros((max(nf, 0), nb))
    for i in range(nf):
        for jj in range(nb):
            for kk in range(nk):
                if res.omega[jj, kk] == freqs[i]:
                    spec[i, jj] += res.mode_kappa[jj, kk]
    with open(os.path.join(directory, "phononconductivity.out"), "w") as f:
        f.write("# Datatype = phononconductivity\n")
        f.write("# Nb = %6d\n" % nf)
        f.write("# Nbranch = %4d\n" % nb)
        f.write("# omega, fluxphononout(1:Nbranch)\n")
        for i in range(nf):
            f.write(fortran_e(freqs[i], 20, 10) + "".join(fortran_e(spec[i, j], 20, 10)
                                                          for j in range(nb)) + "\n")


# ------------------------------------------------------------------ alias-mode estimator

def kappa_bulk(omega, vel, mfp_nm, weight, temperature: float, vol: float) -> float:
    """Bulk conductivity sum_m C_m w_m v_m Lambda_m / (3 V N_q), the kappa of fileoutput
    (RT:2747-2755) with EMFP = bulk MFP; N_q = int(sum(w)/nb) (RT:2657)."""
    omega = np.asarray(omega, dtype=np.float64)
    nb = omega.shape[0]
    nq = int(np.asarray(weight, dtype=np.float64).sum() / nb)
    cap = np.vectorize(lambda o: heat_capacity(float(o), temperature))(omega)
    return float((cap * weight * np.asarray(mfp_nm) * 1.0e-9 * vel).sum() / 3 / vol / nq)


@dataclasses.dataclass
class AliasKappa:
    ratio: float            # kappa / kappa_bulk = E[X], X = (3L/4) T / Lambda_m
    stderr: float           # standard error of ratio
    paths: int
    var_x: float            # per-path variance of X (for the paths needed at a target error)

    def paths_for_rel_error(self, rel: float) -> float:
        return self.var_x / (rel * self.ratio) ** 2 if self.ratio > 0 else math.inf


def alias_kappa(mode_paths, mode_trans, mfp_nm, length: float) -> AliasKappa:
    """kappa/kappa_bulk from alias-mode tallies.  A path of mode m that transmits scores
    X = (3L/4)/Lambda_m, else 0: with modes drawn in proportion to C v Lambda w and the
    Lambertian polar angle, E[X] = sum_m q_m Lambda_eff,m / Lambda_m with
    Lambda_eff = (3L/2) int T sin cos dtheta (RT:2418-2451 in the continuum limit)."""
    n = int(np.sum(mode_paths))
    x = 0.75 * length / np.asarray(mfp_nm, dtype=np.float64)
    t = np.asarray(mode_trans, dtype=np.float64)
    s1 = float((t * x).sum())
    s2 = float((t * x * x).sum())
    mean = s1 / n
    var = max(s2 / n - mean * mean, 0.0)
    return AliasKappa(ratio=mean, stderr=math.sqrt(var / n), paths=n, var_x=var)


@dataclasses.dataclass
class AliasGradient:
    ratio: float              # kappa / kappa_bulk = (3L/4) E[x], x = T / Lambda_m
    stderr: float
    grad: np.ndarray          # d ratio / d theta: ngroup specularities, then nb branch scales
    grad_stderr: np.ndarray
    paths: int
    ngroup: int

    @property
    def grad_p(self):
        return self.grad[:self.ngroup]

    @property
    def grad_s(self):
        return self.grad[self.ngroup:]


def _ratio_of_sums(num, den):
    """sum(num)/sum(den) over batches and its batch-means standard error (ratio estimator:
    residuals num_j - r den_j, variance B/(B-1) sum r_j^2 / (sum den)^2)."""
    num = np.asarray(num, dtype=np.float64)
    den = np.asarray(den, dtype=np.float64)
    tot = den.sum()
    r = num.sum(axis=0) / tot
    res = num - np.multiply.outer(den, r) if num.ndim > 1 else num - den * r
    b = den.size
    var = (res * res).sum(axis=0) * b / max(b - 1, 1) / tot ** 2
    return r, np.sqrt(var)


def alias_gradients(batch_sums, length: float, ngroup: int) -> AliasGradient:
    """kappa/kappa_bulk and its likelihood-ratio gradient from pt_adjoint_alias batch sums
    [n, sum x, sum x^2, sum (x - b) score_k] (x = T / Lambda_m): ratio = (3L/4) E[x],
    d ratio / d theta_k = (3L/4) E[(x - b) score_k], errors by batch means."""
    s = np.asarray(batch_sums, dtype=np.float64)
    n = s[:, 0]
    c = 0.75 * length
    ratio, se = _ratio_of_sums(s[:, 1], n)
    grad, gse = _ratio_of_sums(s[:, 3:], n)
    return AliasGradient(ratio=float(c * ratio), stderr=float(c * se), grad=c * grad,
                         grad_stderr=c * gse, paths=int(n.sum()), ngroup=int(ngroup))


def grain_kappa_shares(grain_tally, npaths: int):
    """Per-grain shares of kappa/kappa_bulk from pt_trace_alias_grains tallies: a path's
    x displacements add up to L T, so kappa/kappa_bulk = (3L/4) E[T/Lambda] =
    (3/4) sum_g tally[g, 0] / npaths; share_g = (3/4) tally[g, 0] / npaths."""
    t = np.asarray(grain_tally, dtype=np.float64)
    return 0.75 * t[:, 0] / float(npaths)


# ---------------------------------------------------------------------------- trace.out

def _flang_shortest(f: np.float32):
    """Shortest decimal digits of a positive REAL(4) the way the flang runtime picks them:
    with lo, hi the midpoints to the neighbouring REAL(4) values, the common leading digits
    of lo and hi, then at the first digit where they differ (dl < dh) the digit
    (dl + dh + 1) // 2.  Returns (digit string, decimal exponent e of d.ddd x 10^e)."""
    from decimal import Decimal, ROUND_FLOOR, localcontext
    with localcontext() as ctx:
        ctx.prec = 80
        x = Decimal(float(f))
        lo = (x + Decimal(float(np.nextafter(f, np.float32(0))))) / 2
        hi = (x + Decimal(float(np.nextafter(f, np.float32(np.inf))))) / 2
        top = hi.adjusted()
        for k in range(1, 40):
            q = Decimal(10) ** (top - k + 1)
            a = int((lo / q).to_integral_value(rounding=ROUND_FLOOR))
            b = int((hi / q).to_integral_value(rounding=ROUND_FLOOR))
            if a != b:
                n = b - b % 10 + (a % 10 + b % 10 + 1) // 2
                digits = str(n).rstrip("0") or "0"
                return digits, top - k + len(str(n))
    raise ValueError("no shortest form for %r" % f)


def fortran_list_real4(x: float) -> str:
    """One REAL(4) value as the flang runtime's list-directed output writes it
    (write(fd,*) real(Position), RT:1663, 1820): the shortest digits that read back to the
    same REAL(4) (chosen as _flang_shortest does); F editing when the value rounded to one
    digit is 0.d x 10^e with 0 <= e <= 6, else E editing with a two-digit exponent
    ("54.351204", ".5", "100.", "0.", "6.893895E-02", "1.E+07")."""
    f = np.float32(x)
    if f == 0:
        return "-0." if np.signbit(f) else "0."
    sign = "-" if f < 0 else ""
    digits, e = _flang_shortest(abs(f))            # value = d.ddd x 10^e
    e1 = int(("%.0e" % abs(float(f))).split("e")[1]) + 1    # one-digit rounding, 0.d x 10^e1
    if 0 <= e1 <= 6:
        k = e + 1                                  # digits before the point
        if k <= 0:
            s = "." + "0" * (-k) + digits
        elif k >= len(digits):
            s = digits + "0" * (k - len(digits)) + "."
        else:
            s = digits[:k] + "." + digits[k:]
    else:
        s = digits[0] + "." + digits[1:] + "E" + ("-" if e < 0 else "+") + "%02d" % abs(e)
    return sign + s


def trace_out_lines(trajectories, particle_trace: Optional[int] = None,
                    interval_trace: Optional[int] = None, fp64: bool = False):
    """Lines of trace.out (RT:3377-3382 header, then per traced particle RT:1661-1665,
    1818-1831): the start position and the position after every event, one per line, and
    two empty lines after each particle.  fp64=True writes full FP64 positions ("%.17g")
    instead of the reference's REAL(4) list-directed values.  The header is written when
    particle_trace and interval_trace are given."""
    out = []
    if particle_trace is not None and interval_trace is not None:
        out += ["# Datatype = trace", "# particle_trace = %8d" % particle_trace,
                "# interval_trace = %8d" % interval_trace, ""]
    for t in trajectories:
        for p in np.asarray(t, dtype=np.float64).reshape(-1, 3):
            if fp64:
                out.append(" " + " ".join("%.17g" % v for v in p))
            else:
                out.append(" " + " ".join(fortran_list_real4(v) for v in p))
        out += ["", ""]
    return out


def write_trace(path: str, trajectories, particle_trace: Optional[int] = None,
                interval_trace: Optional[int] = None, fp64: bool = False):
    with open(path, "w") as f:
        for ln in trace_out_lines(trajectories, particle_trace, interval_trace, fp64):
            f.write(ln + "\n")
