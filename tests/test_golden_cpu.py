"""Pins the FP64 oracle (oracle/nb6_oracle.c) against the golden vectors that the
reference's own AVX back end produced (tests/golden/make_golden.py).

The reference works in FP32 absolute coordinates (reg.avx.cpp:149-152,
212-264), so its neighbour decision is uncertain for pairs whose min(r2, r2p)
lies within its own rounding band of the threshold; oracle/refband.py derives
that band per pair from the FP32 operations the reference performs.  Every
entry where the oracle's lists and the reference's differ must lie inside it
(and the off-centre cases show that some lie outside one FP32 ulp of the
threshold, where only the reference's rounding decides).  Forces must agree
to the reference's own FP32 accuracy.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import refband as RB

GOLD = Path(__file__).resolve().parent / "golden"


def ref_rows(g, lmax):
    """The golden (counts, entries) as [ni][lmax] rows."""
    rows = np.zeros((len(g["counts"]), lmax), np.int32)
    k = 0
    for i, c in enumerate(g["counts"]):
        rows[i, 0] = c
        if c > 0:
            rows[i, 1:1 + c] = g["entries"][k:k + c]
            k += c
    return rows


def in_band(g, i):
    """Number of j whose FP64 min(r2, r2p) lies within the reference's rounding band
    of row i's threshold."""
    rmin, thr, band = RB.pair_geometry(g["xi"][i], g["vi"][i], g["dtr"][i], g["h2"][i], g["x"],
                                       g["v"], g["m"], int(g["m_flag"]))
    return int((np.abs(rmin - thr) <= band).sum())


def rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


@pytest.mark.parametrize("case,acc_tol,jrk_tol", [
    ("reg_plummer1k", 5e-6, 2e-5),
    ("reg_kroupa16k_m0", 5e-6, 5e-5),
    ("reg_kroupa16k_m1", 5e-6, 5e-5),
])
def test_oracle_regf_matches_reference_avx(case, acc_tol, jrk_tol):
    g = dict(np.load(GOLD / f"{case}.npz"))
    lmax, nnbmax, mf = int(g["lmax"]), int(g["nnbmax"]), int(g["m_flag"])
    acc, jrk, pot, lst = O.regf(g["m"], g["x"], g["v"], g["h2"], g["dtr"], g["xi"], g["vi"],
                                lmax, nnbmax, mf)
    rr = ref_rows(g, lmax)
    c = RB.classify_rows(lst, rr, g["x"], g["v"], g["m"], g["xi"], g["vi"], g["h2"], g["dtr"], mf)
    print(case, c)
    assert c["overflow_sign_diff"] == 0 and c["outside_ref_band"] == 0
    assert c["rows_identical"] >= 0.9 * c["rows"]
    clean = np.array([i for i in range(len(lst)) if lst[i, 0] >= 0 and
                      np.array_equal(lst[i, :1 + lst[i, 0]], rr[i, :1 + rr[i, 0]])])
    ea = rel(acc[clean], g["acc"][clean])
    ej = rel(jrk[clean], g["jrk"][clean])
    ep = np.abs(pot[clean] - g["pot"][clean]) / g["pot"][clean]
    print(f"{case}: ref-vs-fp64 acc max {ea.max():.2e} jrk max {ej.max():.2e} pot max {ep.max():.2e}")
    assert ea.max() < acc_tol and ep.max() < acc_tol and ej.max() < jrk_tol


@pytest.mark.parametrize("case", ["reg_shifted16k", "reg_kingbin16k"])
def test_oracle_lists_differ_from_reference_only_inside_its_band(case):
    """Off-centre clusters (|x| ~ 4-40): the reference's FP32 absolute coordinates
    flip decisions for pairs that are more than one FP32 ulp from the threshold, but
    never outside its own rounding band; its forces carry the FP32 error of absolute
    coordinates (median still ~1e-6, a dropped term or a sign would not be)."""
    g = dict(np.load(GOLD / f"{case}.npz"))
    lmax, nnbmax, mf = int(g["lmax"]), int(g["nnbmax"]), int(g["m_flag"])
    acc, jrk, pot, lst = O.regf(g["m"], g["x"], g["v"], g["h2"], g["dtr"], g["xi"], g["vi"],
                                lmax, nnbmax, mf)
    c = RB.classify_rows(lst, ref_rows(g, lmax), g["x"], g["v"], g["m"], g["xi"], g["vi"],
                         g["h2"], g["dtr"], mf)
    print(case, c)
    assert c["overflow_sign_diff"] == 0 and c["outside_ref_band"] == 0
    assert c["diff_entries"] > 0 and c["outside_ulp"] > 0      # the case exercises the band
    ea = rel(acc, g["acc"])
    ep = np.abs(pot - g["pot"]) / g["pot"]
    print(f"{case}: ref-vs-fp64 acc median {np.median(ea):.2e} max {ea.max():.2e}")
    assert np.median(ea) < 5e-6 and np.median(ep) < 5e-6


def test_reference_band_is_not_vacuous():
    """A mistake planted in the oracle's lists (a neighbour dropped in a row, a
    non-neighbour added) is caught: those pairs lie far outside the band."""
    g = dict(np.load(GOLD / "reg_kroupa16k_m0.npz"))
    lmax, nnbmax = int(g["lmax"]), int(g["nnbmax"])
    _, _, _, lst = O.regf(g["m"], g["x"], g["v"], g["h2"], g["dtr"], g["xi"], g["vi"], lmax,
                          nnbmax, 0)
    bad = lst.copy()
    k = int(np.argmax(bad[:, 0] > 10))
    c0 = bad[k, 0]
    bad[k, 1:c0] = bad[k, 2:c0 + 1]                     # drop the first neighbour
    bad[k, 0] = c0 - 1
    c = RB.classify_rows(bad, ref_rows(g, lmax), g["x"], g["v"], g["m"], g["xi"], g["vi"],
                         g["h2"], g["dtr"], 0)
    assert c["outside_ref_band"] == 1 and c["outside_ulp"] == 1


def test_oracle_regf_overflow_matches_reference_avx():
    """-count for rows above nnbmax, with the exact count (reg.avx.cpp:320-322)."""
    g = dict(np.load(GOLD / "reg_overflow4k.npz"))
    acc, jrk, pot, lst = O.regf(g["m"], g["x"], g["v"], g["h2"], g["dtr"], g["xi"], g["vi"],
                                int(g["lmax"]), int(g["nnbmax"]), 0)
    assert (g["counts"] < 0).any() and (g["counts"] > 0).any()
    for i, c in enumerate(g["counts"]):
        amb = in_band(g, i)
        assert abs(abs(int(lst[i, 0])) - abs(int(c))) <= amb
        if abs(abs(int(c)) - int(g["nnbmax"])) > amb:
            assert (lst[i, 0] < 0) == (c < 0)


def test_oracle_pot_matches_reference_avx():
    """The AVX potential splits positions at 2^-16 (pot.avx.cpp:46-54): the FP32 low
    part resolves ~1e-12, so a binary at 0.005 AU (2.4e-8 N-body units) carries
    ~5e-5 relative error in the reference itself; elsewhere it agrees to 1e-6."""
    g = dict(np.load(GOLD / "pot_binaries4k.npz"))
    p = O.pot(g["m"], g["x"])
    err = np.abs(p - g["pot"]) / g["pot"]
    print(f"pot ref-vs-fp64 max {err.max():.2e} median {np.median(err):.2e}")
    assert err.max() < 1e-4 and np.median(err) < 1e-6


def test_oracle_firr_matches_reference_avx():
    g = dict(np.load(GOLD / "irr_plummer1k.npz"))
    acc, jrk, nn = O.firr(float(g["ti"]), g["addr"], g["x"], g["v"], g["a2"], g["j6"], g["m"],
                          g["t0"], g["lists"], 400)
    ea, ej = rel(acc, g["acc"]), rel(jrk, g["jrk"])
    print(f"irr ref-vs-fp64 acc max {ea.max():.2e} jrk max {ej.max():.2e}")
    assert ea.max() < 1e-5 and ej.max() < 1e-4
    # nearest neighbour: equal, or an FP32 near-tie of the reference
    x, v, a2, j6, t0 = g["x"], g["v"], g["a2"], g["j6"], g["t0"]
    ti = float(g["ti"])

    def pred(k):
        s = ti - t0[k]
        return x[k] + s * (v[k] + s * (a2[k] + s * j6[k]))
    for k, a in enumerate(g["addr"]):
        if nn[k] != g["nnbid"][k]:
            xi = pred(a - 1)
            r_mine = ((pred(nn[k] - 1) - xi) ** 2).sum()
            r_ref = ((pred(g["nnbid"][k] - 1) - xi) ** 2).sum()
            assert abs(r_mine - r_ref) <= 1e-5 * r_mine
