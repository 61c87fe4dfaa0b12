"""Checker (test infrastructure, never product code): how the neighbour lists and
forces of this library relate to the reference's own FP32 back ends.

The reference rounds absolute FP64 positions and velocities to FP32 and forms the
neighbour criterion in FP32:

    dx  = xj - xi,  dv = vj - vi                       reg.avx.cpp:149-152, 212-246
    dxp = dx + dtr dv                                  reg.avx.cpp:248-250
    r2  = |dx|^2,  r2p = |dxp|^2,  rmin = min(r2, r2p) reg.avx.cpp:252-256
    neighbour iff rmin < h2   (or < m_j h2)            reg.avx.cpp:257-264
    (the CUDA kernel is the same FP32 recipe:          gpunb.velocity.cu:64-72, 169-190)

so its decision for a pair can differ from the exact (FP64) one only when the
exact rmin lies within the reference's own rounding error of the threshold.
``ref_band`` bounds that error per pair from first principles (u = 2^-24, round to
nearest, with or without FMA contraction):

    e_dx[c]  <= u (|xi[c]| + |xj[c]| + |dx[c]|)        two input roundings + the difference
    e_dv[c]  <= u (|vi[c]| + |vj[c]| + |dv[c]|)
    e_dxp[c] <= e_dx[c] + dtr e_dv[c] + 2u dtr |dv[c]| + u |dxp[c]|
    |r2_ref - r2|   <= sum_c (2 |dx[c]| e_dx[c] + e_dx[c]^2) + 3u r2     (same for r2p)
    |thr_ref - thr| <= u h2   (m_flag: 3u m_j h2)

and ``SAFETY`` (1.5) covers the (1 + u) factors dropped above.  ``classify_rows``
compares two sets of list rows pair by pair and counts the differing entries that lie
outside one FP32 ulp of the threshold and outside the reference band.  A pair of this
library's lists (which equal the FP64 oracle's) that differs from the reference's
outside the band would be a real disagreement; inside it, the reference's own
rounding decides.
"""
from __future__ import annotations

import numpy as np

U = 2.0 ** -24
SAFETY = 1.5


def ulp32(t):
    """Spacing of float32 numbers at |t| (as float64)."""
    return np.spacing(np.abs(np.asarray(t, np.float64)).astype(np.float32)).astype(np.float64)


def pair_geometry(xi, vi, dtr, h2, xj, vj, mj, m_flag):
    """FP64 rmin = min(r2, r2p), threshold and the reference's error bound for the
    pairs (i, j) given by the rows of xj, vj, mj."""
    dx = xj - xi
    dv = vj - vi
    dxp = dx + dtr * dv
    r2 = (dx * dx).sum(-1)
    r2p = (dxp * dxp).sum(-1)
    rmin = np.minimum(r2, r2p)
    e_dx = U * (np.abs(xi) + np.abs(xj) + np.abs(dx))
    e_dv = U * (np.abs(vi) + np.abs(vj) + np.abs(dv))
    e_dxp = e_dx + dtr * e_dv + 2.0 * U * dtr * np.abs(dv) + U * np.abs(dxp)
    d_r2 = (2.0 * np.abs(dx) * e_dx + e_dx * e_dx).sum(-1) + 3.0 * U * r2
    d_r2p = (2.0 * np.abs(dxp) * e_dxp + e_dxp * e_dxp).sum(-1) + 3.0 * U * r2p
    if m_flag:
        thr = mj * h2
        d_thr = 3.0 * U * thr
    else:
        thr = np.full_like(rmin, h2)
        d_thr = U * thr
    band = SAFETY * (np.maximum(d_r2, d_r2p) + d_thr)
    return rmin, thr, band


def _row_set(row):
    c = int(row[0])
    return set(row[1:1 + c].tolist()) if c >= 0 else None


def classify_rows(rows_a, rows_b, x, v, m, xi, vi, h2, dtr, m_flag, index_base=0):
    """Compare list rows a and b ([ni][lmax], row format [count, j...], j 0-based into
    x, v, m after subtracting index_base).  Returns counts:

    rows                 rows compared (both non-negative counts)
    rows_identical       rows whose sets are equal
    overflow_sign_diff   rows where exactly one side overflowed
    diff_entries         entries in one row and not the other
    outside_ulp          of those, |rmin - thr| > one FP32 ulp of thr
    outside_ref_band     of those, |rmin - thr| > the reference's rounding band
    max_band_ratio       max |rmin - thr| / band over the differing entries
    """
    out = dict(rows=0, rows_identical=0, overflow_sign_diff=0, diff_entries=0,
               outside_ulp=0, outside_ref_band=0, max_band_ratio=0.0)
    for i in range(len(rows_a)):
        sa, sb = _row_set(rows_a[i]), _row_set(rows_b[i])
        if (sa is None) != (sb is None):
            out["overflow_sign_diff"] += 1
            continue
        if sa is None:
            continue
        out["rows"] += 1
        d = np.array(sorted(sa ^ sb), dtype=np.int64) - index_base
        if d.size == 0:
            out["rows_identical"] += 1
            continue
        rmin, thr, band = pair_geometry(xi[i], vi[i], dtr[i], h2[i], x[d], v[d], m[d], m_flag)
        gap = np.abs(rmin - thr)
        out["diff_entries"] += int(d.size)
        out["outside_ulp"] += int((gap > ulp32(thr)).sum())
        out["outside_ref_band"] += int((gap > band).sum())
        out["max_band_ratio"] = max(out["max_band_ratio"], float((gap / band).max()))
    return out


def force_rule(ours, ref, fp64, tol=1e-6):
    """Per-particle rule |ours - ref| <= tol |ref| + |ref - fp64| (vector norms over the
    last axis).  Returns the number of particles that break it and the largest
    |ours - ref| / (tol |ref| + |ref - fp64|)."""
    ours, ref, fp64 = (np.atleast_2d(np.asarray(a, np.float64)) for a in (ours, ref, fp64))
    if ours.shape[0] == 1 and ours.shape[1] != 1:
        ours, ref, fp64 = ours.T, ref.T, fp64.T
    lhs = np.linalg.norm(ours - ref, axis=1)
    rhs = tol * np.linalg.norm(ref, axis=1) + np.linalg.norm(ref - fp64, axis=1)
    ratio = np.where(rhs > 0, lhs / np.where(rhs > 0, rhs, 1.0), np.where(lhs > 0, np.inf, 0.0))
    return int((ratio > 1.0).sum()), float(ratio.max()) if ratio.size else 0.0
