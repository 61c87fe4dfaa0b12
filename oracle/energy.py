"""FP64 restatement of the energy path, the GPU branch of ENERGY (energy.F:28-62,
:119-124) with PHICOR (phicor.f:1-76) and the EBIN sum (energy.F:13-24) -- TEST
INFRASTRUCTURE ONLY: imported by tests/ as the checker of nb6_energy_device.

Bodies are numbered 1..NTOT as in common6.h: KS components 1..2 NPAIRS, singles
IFIRST = 2 NPAIRS + 1 .. N, centres of mass N+1 .. NTOT = N + NPAIRS.  The potential
that GPUPOT returns is taken here from an FP64 all-pairs sum over IFIRST..NTOT
(phi_i = sum_{j, r_ij > 0} m_j / r_ij, positive).  Arrays: x, xdot [nmax][3], body
[nmax], lst int [nmax][lmax] (count, then 1-based indices), ht [npairs].
"""
from __future__ import annotations

import math

import numpy as np


def gpupot_fp64(x, body, i0, i1):
    """phi of bodies i0..i1 (1-based, inclusive) from the same set, r > 0 only."""
    xs = np.asarray(x[i0 - 1:i1], float)
    ms = np.asarray(body[i0 - 1:i1], float)
    out = np.empty(len(ms))
    for k in range(len(ms)):
        d = xs - xs[k]
        r = np.sqrt((d * d).sum(1))
        sel = r > 0
        out[k] = (ms[sel] / r[sel]).sum()
    return out


def _d2(x, a, b):
    """|x_a - x_b|^2 with 1-based a, b in the order of phicor.f:18-20."""
    r2 = 0.0
    for k in range(3):
        d = x[a - 1][k] - x[b - 1][k]
        r2 = r2 + d * d
    return r2


def phicor(I, x, body, lst, n, thr):
    """(DPHI1, DPHI2) of phicor.f for body I (1-based)."""
    d1 = d2 = 0.0
    if I <= n:
        nnb = int(lst[I - 1][0])
        for L in range(nnb, 0, -1):
            J = int(lst[I - 1][L])
            if J <= n:
                break
            rij2 = _d2(x, I, J)
            if rij2 < thr:
                j1 = 2 * (J - n) - 1
                j2 = j1 + 1
                rj1 = _d2(x, I, j1)
                rj2 = _d2(x, I, j2)
                d1 = d1 - body[J - 1] / math.sqrt(rij2) + body[j1 - 1] / math.sqrt(rj1) \
                    + body[j2 - 1] / math.sqrt(rj2)
    else:
        i1 = 2 * (I - n) - 1
        i2 = i1 + 1
        npt = int(lst[i1 - 1][0])
        for L in range(1, npt + 1):
            J = int(lst[i1 - 1][L])
            rij2 = _d2(x, I, J)
            if J <= n or rij2 > thr:
                ri1 = _d2(x, i1, J)
                ri2 = _d2(x, i2, J)
                d1 = d1 - body[J - 1] / math.sqrt(rij2) + body[J - 1] / math.sqrt(ri1)
                d2 = d2 - body[J - 1] / math.sqrt(rij2) + body[J - 1] / math.sqrt(ri2)
            else:
                d1 = d1 - body[J - 1] / math.sqrt(rij2)
                d2 = d2 - body[J - 1] / math.sqrt(rij2)
                j1 = 2 * (J - n) - 1
                for jj in (j1, j1 + 1):
                    rj1 = _d2(x, i1, jj)
                    rj2 = _d2(x, i2, jj)
                    d1 = d1 + body[jj - 1] / math.sqrt(rj1)
                    d2 = d2 + body[jj - 1] / math.sqrt(rj2)
    return d1, d2


def energy(x, xdot, body, lst, n, npairs, thr, ht=None, phi=None):
    """Returns (ZKIN, POT, EBIN, PHIDBL[nmax]).  phi: GPUPOT's output for IFIRST..NTOT
    (default: the FP64 all-pairs sum)."""
    x = np.asarray(x, float)
    body = np.asarray(body, float)
    ifirst = 2 * npairs + 1
    ntot = n + npairs
    if phi is None:
        phi = gpupot_fp64(x, body, ifirst, ntot)
    phidbl = np.zeros(len(body))
    phidbl[ifirst - 1:ntot] = phi
    for p in range(1, npairs + 1):                       # energy.F:36-44
        phidbl[2 * p - 2] = phidbl[n + p - 1]
        phidbl[2 * p - 1] = phidbl[n + p - 1]
    pot = 0.0
    for I in range(ifirst, ntot + 1):                     # energy.F:47-58
        d1, d2 = phicor(I, x, body, lst, n, thr)
        if I <= n:
            phidbl[I - 1] = phidbl[I - 1] + d1
            pot = pot + body[I - 1] * phidbl[I - 1]
        else:
            i1 = 2 * (I - n) - 1
            phidbl[i1 - 1] = phidbl[i1 - 1] + d1
            phidbl[i1] = phidbl[i1] + d2
            pot = pot + body[i1 - 1] * phidbl[i1 - 1] + body[i1] * phidbl[i1]
    pot = 0.5 * pot
    zkin = 0.0
    for I in range(ifirst, ntot + 1):                     # energy.F:119-124
        v = xdot[I - 1]
        zkin = zkin + body[I - 1] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    zkin = 0.5 * zkin
    ebin = 0.0
    for p in range(1, npairs + 1):                        # energy.F:13-24
        mcm = body[n + p - 1]
        if mcm > 0.0:
            ebin = ebin + body[2 * p - 2] * body[2 * p - 1] * ht[p - 1] / mcm
    return zkin, pot, ebin, phidbl
