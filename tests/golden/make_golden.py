#!/usr/bin/env python3
"""Writes the golden vectors in this directory by running the reference's own
CPU back end -- src/Main/reg.avx.cpp, pot.avx.cpp and irr.avx.cpp, compiled from
the read-only reference tree into oracle/_ref by ``make -C oracle ref`` -- on
seeded synthetic inputs.  Inputs and outputs are stored (npz, compressed), so
the tests never need the reference tree or the generators to reproduce them.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path[:0] = [str(ROOT / "oracle"), str(ROOT / "nbody6ppgpu-beijing_amd"), str(ROOT / "tests")]

import oracle as O  # noqa: E402
from helpers import c1_inputs, c2_inputs  # noqa: E402
from nb6 import ics  # noqa: E402


def pack_lists(lst):
    """Row-format lists -> (counts, concatenated entries)."""
    counts = lst[:, 0].astype(np.int32)
    ent = [lst[i, 1:1 + c] for i, c in enumerate(counts) if c > 0]
    return counts, (np.concatenate(ent).astype(np.int32) if ent else np.zeros(0, np.int32))


def reg_case(name, m, x, v, h2, dtr, xi, vi, lmax, nnbmax, m_flag):
    R = O.RefReg(m.size)
    R.send(m, x, v)
    acc, jrk, pot, lst = R.regf(h2, dtr, xi, vi, lmax, nnbmax, m_flag)
    R.close()
    counts, ent = pack_lists(lst)
    np.savez_compressed(HERE / f"{name}.npz", m=m, x=x, v=v, h2=h2, dtr=dtr, xi=xi, vi=vi,
                        lmax=lmax, nnbmax=nnbmax, m_flag=m_flag, acc=acc, jrk=jrk, pot=pot,
                        counts=counts, entries=ent)
    print(f"{name}: ni={h2.size} nj={m.size} nnb mean {counts[counts > 0].mean():.1f}")


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    # config 1: equal-mass Plummer N=1024, all i, ~50 neighbours
    m, x, v, h2, dtr = c1_inputs()
    reg_case("reg_plummer1k", m, x, v, h2, dtr, x, v, 400, 350, 0)
    # config-2-like at N=16384: Kroupa masses, random i-subset, both criteria
    for mf in (0, 1):
        m, x, v, idx, h2, dtr = c2_inputs(n=16384, ni=256, m_flag=mf, seed=20261210 + mf)
        reg_case(f"reg_kroupa16k_m{mf}", m, x, v, h2, dtr, x[idx], v[idx], 500, 450, mf)
    # overflow: nnbmax below the counts of part of the rows
    m, x, v, idx, h2, dtr = c2_inputs(n=4096, ni=128, counts_median=60.0, seed=20261212)
    reg_case("reg_overflow4k", m, x, v, h2, dtr, x[idx], v[idx], 120, 64, 0)
    # off-centre clusters: the reference's FP32 absolute coordinates then flip neighbour
    # decisions near the threshold (pairs inside its rounding band)
    m, x, v, idx, h2, dtr = c2_inputs(n=16384, ni=512, seed=20261214)
    x = x + np.array([30.0, -20.0, 10.0])
    reg_case("reg_shifted16k", m, x, v, h2, dtr, x[idx], v[idx], 500, 450, 0)
    mass = ics.kroupa_masses(16384, 20261215)
    m, x, v = ics.king(16384, 20261216, w0=6.0, masses=mass)
    m, x, v = ics.add_binaries(m, x, v, 0.5, 20261217)
    x = x + np.array([3.0, 2.0, -4.0])
    rng = np.random.Generator(np.random.PCG64(20261218))
    idx = np.sort(rng.choice(m.size, size=512, replace=False))
    rs = ics.rs_for_counts(x, idx, ics.lognormal_counts(512, 20261219, median=80.0))
    dtr = np.exp(rng.uniform(np.log(2.0 ** -10), np.log(2.0 ** -4), 512))
    reg_case("reg_kingbin16k", m, x, v, rs * rs, dtr, x[idx], v[idx], 600, 550, 0)

    # potential: Kroupa Plummer with 50% binaries, N=4096 (all i)
    mm = ics.kroupa_masses(4096, 31)
    m, x, v = ics.plummer(4096, 32, masses=mm)
    m, x, v = ics.add_binaries(m, x, v, 0.5, 33)
    pot = O.ref_pot(m, x)
    np.savez_compressed(HERE / "pot_binaries4k.npz", m=m, x=x, pot=pot)
    print("pot_binaries4k: n=4096")

    # irregular force: N=1024, predictor records at block times, lists from the FP64
    # regular-force neighbour sets (self removed, 1-based), ti one block later
    m, x, v, h2, dtr = c1_inputs(n=1024, seed=41)
    n = m.size
    rng = np.random.Generator(np.random.PCG64(42))
    a2 = 0.5 * rng.normal(0, 0.3, (n, 3))
    j6 = rng.normal(0, 0.2, (n, 3)) / 6.0
    t0 = np.floor(rng.uniform(0, 64, n)) / 1024.0
    _, _, _, nb = O.regf(m, x, v, h2, dtr, x, v, 400, 350)
    lists = np.zeros((n, 400), np.int32)
    for i in range(n):
        row = [j + 1 for j in nb[i, 1:1 + nb[i, 0]] if j != i]
        lists[i, 0] = len(row)
        lists[i, 1:1 + len(row)] = row
    ti = 64.0 / 1024.0
    addr = np.arange(1, n + 1, dtype=np.int32)[::2]
    R = O.RefIrr(n, 400)
    for i in range(n):
        R.set_jp(i + 1, x[i], v[i], a2[i], j6[i], m[i], t0[i])
        R.set_list(i + 1, lists[i])
    acc, jrk, nn = R.firr_vec(ti, addr)
    R.close()
    np.savez_compressed(HERE / "irr_plummer1k.npz", m=m, x=x, v=v, a2=a2, j6=j6, t0=t0,
                        lists=lists, ti=ti, addr=addr, acc=acc, jrk=jrk, nnbid=nn)
    print("irr_plummer1k: n=1024")


if __name__ == "__main__":
    main()
