"""Shared input builders and comparators for the tests (seeded, deterministic)."""
from __future__ import annotations

import numpy as np

from nb6 import ics

LMAX_SMALL = 400     # configure.ac: LMAX for --with-par <= 16k
LMAX_LARGE = 600     # LMAX for --with-par >= 128k


def nnbmax_for(n, lmax):
    """NNBMAX = min(N/2, LMAX-50) (input.F:132)."""
    return min(n // 2, lmax - 50)


def c1_inputs(seed=20261203, n=1024, nnbopt=50):
    """Config-1-like: equal-mass Plummer, RS for ~nnbopt neighbours, initial
    regular step STEPR = min(SMAX*sqrt(1+r^2), SMAX) with SMAX = 0.125 (fpoly0.F:57-59)."""
    m, x, v = ics.plummer(n, seed)
    rs = ics.rs_for_counts(x, np.arange(n), np.full(n, nnbopt))
    r2 = (x * x).sum(1)
    smax = 0.125
    dtr = np.minimum(smax * np.sqrt(1.0 + r2), smax) * 0.05
    return m, x, v, rs * rs, dtr


def c2_inputs(seed=20261204, n=65536, ni=4096, m_flag=0, counts_median=80.0):
    """Config-2-like: Kroupa Plummer, random i-subset, log-normal neighbour counts
    (median 80, cap 500), dtr log-uniform in [2^-10, 2^-4]."""
    mass = ics.kroupa_masses(n, seed)
    m, x, v = ics.plummer(n, seed + 1, masses=mass)
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    idx = np.sort(rng.choice(n, size=ni, replace=False))
    counts = ics.lognormal_counts(ni, seed + 3, median=counts_median)
    rs = ics.rs_for_counts(x, idx, counts)
    h2 = rs * rs
    if m_flag:
        h2 = h2 / m.mean()      # RS^2 / BODYM (util_gpu.F:46-48)
    dtr = np.exp(rng.uniform(np.log(2.0 ** -10), np.log(2.0 ** -4), ni))
    return m, x, v, idx, h2, dtr


def vec_rel_err(a, ref):
    """Per-particle |a - ref| / |ref| (vector norm); 0/0 -> 0."""
    num = np.linalg.norm(np.atleast_2d(a - ref).reshape(len(ref), -1), axis=1)
    den = np.linalg.norm(np.atleast_2d(ref).reshape(len(ref), -1), axis=1)
    out = np.zeros_like(num)
    nz = den > 0
    out[nz] = num[nz] / den[nz]
    out[~nz] = num[~nz]
    return out


def rows_as_sets(lst):
    out = []
    for row in lst:
        c = int(row[0])
        out.append(tuple(row[1:1 + c]) if c >= 0 else ("ovf", -c))
    return out
