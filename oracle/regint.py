"""FP64 restatement of the reference's CPU regular force of the build without a
GPU back end, regint.f:1-102 (TEST INFRASTRUCTURE ONLY: imported by tests/).

Per active particle I (regint.f:25-97):
  * every J != I of [IFIRST, N] is tested: RSM = min(r^2, |dx + dtr dv|^2), divided by
    m_J when M_FLAG = 1, against RS2 = RS(I)^2 (/ BODYM when M_FLAG = 1);
  * RSM > RS2: the pair adds m_J dx / r^3 to FREG and (dv - 3 (dx.dv)/r^2 dx) m_J / r^3
    to FDR; otherwise J is a neighbour (the count only; entries kept up to LMAX-3);
  * POT(I) += m_J / r for every J != I -- neighbours included, and added to the value
    POT(I) already holds (regint.f:78);
  * NNB - 1 > LMAX - 4 neighbours (regint.f:81): RS(I) <- NNBOPT / NNB**0.3333 * RS(I)
    when NB_FLAG = 1 (Fortran precedence: NNB**0.3333 is taken first, so RS grows
    whenever NNB < NNBOPT**3) or 0.9 RS(I) when NB_FLAG = 0, and the particle is
    evaluated again from the start (GO TO 2), its POT included.

Differences from gpunb_regf_: the test is <= (regf: <), self is not listed, the
potential includes the neighbours and accumulates, and overflow is resolved inside.
"""
from __future__ import annotations

import numpy as np


class NoConvergence(RuntimeError):
    pass


def regint(x, xdot, body, rs, dtr, ireg, lmax, nnbopt, *, bodym=1.0, m_flag=0, nb_flag=1,
           pot=None, max_retry=50):
    """Returns (freg [ni][3], fdr [ni][3], lists [ni][lmax] (0: count, then 0-based J),
    pot [n], rs [n], retries [ni]).  ireg: 0-based indices of the active particles."""
    x = np.asarray(x, float)
    xdot = np.asarray(xdot, float)
    body = np.asarray(body, float)
    rs = np.array(rs, float)
    pot = np.zeros(len(body)) if pot is None else np.array(pot, float)
    ni = len(ireg)
    freg = np.zeros((ni, 3))
    fdr = np.zeros((ni, 3))
    lists = np.zeros((ni, lmax), np.int32)
    retries = np.zeros(ni, np.int64)
    for ii, i in enumerate(ireg):
        for attempt in range(max_retry + 1):
            rs2 = rs[i] * rs[i]
            if m_flag:
                rs2 = rs2 / bodym
            j = np.r_[0:i, i + 1:len(body)]
            a = x[j] - x[i]
            dv = xdot[j] - xdot[i]
            rij2 = (a * a).sum(1)
            dp = a + dv * dtr[i]
            rijp = (dp * dp).sum(1)
            dr2i = 1.0 / rij2
            dr3i = body[j] * dr2i * np.sqrt(dr2i)
            drdv = (a * dv).sum(1)
            drdp = 3.0 * drdv * dr2i
            rsm = np.minimum(rij2, rijp)
            if m_flag:
                rsm = rsm / body[j]
            far = rsm > rs2
            freg[ii] = (a[far] * dr3i[far, None]).sum(0)
            fdr[ii] = ((dv[far] - a[far] * drdp[far, None]) * dr3i[far, None]).sum(0)
            pot[i] += (dr3i * rij2).sum()
            nbs = j[~far]
            nnb = 1 + nbs.size
            if nnb > lmax - 3:
                retries[ii] += 1
                rs[i] = (float(nnbopt) / float(nnb) ** 0.3333 * rs[i]) if nb_flag else 0.9 * rs[i]
                continue
            lists[ii, 0] = nbs.size
            lists[ii, 1:1 + nbs.size] = nbs
            break
        else:
            raise NoConvergence(f"particle {i}: neighbour list still overflowing after "
                                f"{max_retry} retries (RS = {rs[i]:.3e})")
    return freg, fdr, lists, pot, rs, retries
