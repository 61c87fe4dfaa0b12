"""Plummer N = 1024, equal masses, N-body units, NNBOPT = 50, LMAX = 400,
integrated by regular block steps (nb6/rsdrive.py: start-up
as fpoly0.F, block prediction, gpunb_regf with the particle's RS, util_gpu.F
overflow re-runs, regcor_gpu.F RS rule, Hermite corrector).  Every regular call's
lists, acc, jerk and pot are checked against the FP64 oracle on the same inputs."""
import numpy as np
import pytest

import oracle as O
from helpers import rows_as_sets, vec_rel_err
from nb6 import ics
from nb6.rsdrive import Sim

pytestmark = pytest.mark.gpu


def test_c1_every_regular_call_with_bounded_energy_error():
    """Two N-body time units of the N = 1024 Plummer sphere with a minimum step small
    enough (2^-24) that close encounters are resolved without KS regularisation:
    the relative energy error stays below 1e-4 (measured: 4e-5 at t = 1 and 2, and
    4e-5 at t = 5.4 in a longer run), and every regular call along the way is
    checked against the FP64 oracle on its own inputs (lists identical; acc, jerk and
    pot within 1e-6 for every particle)."""
    n = 1024
    m, x, v = ics.plummer(n, 20261202)
    s = Sim(m, x, v, nnbopt=50, lmax=400, smax=2.0 ** -3, dtmin=2.0 ** -24, eta=0.03)
    worst = {"calls": 0, "particles": 0, "pot": 0.0, "acc": [], "jrk": [], "lists_equal": True}

    def hook(jx, jv, xi, vi, h2, dtr, out, rows):
        racc, rjrk, rpot, rlst = O.regf(m, jx, jv, h2, dtr, xi, vi, s.lmax, s.nnbmax)
        worst["calls"] += 1
        worst["particles"] += h2.size
        worst["lists_equal"] &= rows_as_sets(rows) == rows_as_sets(rlst)
        worst["pot"] = max(worst["pot"], float((np.abs(out[6] - rpot) / rpot).max()))
        worst["acc"].append(vec_rel_err(out[0:3].T, racc))
        worst["jrk"].append(vec_rel_err(out[3:6].T, rjrk))

    try:
        s.start()
        e0 = s.energy()
        s.hook = hook
        while s.t < 1.0:
            s.block_step("rule")
        e1 = s.energy()
        while s.t < 2.0:
            s.block_step("rule")
        s.hook = None
        e2 = s.energy()
    finally:
        s.close()
    ea = np.concatenate(worst.pop("acc"))
    ej = np.concatenate(worst.pop("jrk"))
    d1, d2 = abs(e1 - e0) / abs(e0), abs(e2 - e0) / abs(e0)
    print(f"C1: {worst['calls']} regular calls, {worst['particles']} particle-steps, "
          f"{s.stats['overflow_events']} overflow re-runs; acc max {ea.max():.2e} "
          f"p99.9 {np.percentile(ea, 99.9):.2e}; jerk max {ej.max():.2e} p99 "
          f"{np.percentile(ej, 99):.2e}; pot max {worst['pot']:.2e}; energy {e0:.7f} -> "
          f"{e1:.7f} (t=1, dE/E {d1:.1e}) -> {e2:.7f} (t=2, dE/E {d2:.1e})", flush=True)
    assert worst["lists_equal"]
    assert ea.max() <= 1e-6 and worst["pot"] <= 1e-6 and ej.max() <= 1e-6
    assert abs(e0 + 0.25) < 0.015                # N-body units at the start
    assert d1 < 1e-4 and d2 < 1e-4


def test_run_from_the_reference_example_input_file():
    """examples/example_1k/1k.inp (N = 1000, NNBOPT = 80, ETAR = 0.01, RS0 = 0.15,
    SMAX = 1, DTMIN = 2.5e-6, KZ(22) = 0) read by nb6.io, set up by input_setup and
    advanced to t = 0.25 with every regular call checked against the FP64 oracle."""
    import time
    from pathlib import Path
    from nb6.rsdrive import input_setup
    inp = Path(__file__).resolve().parent / "golden" / "inputs" / "1k.inp"
    m, x, v, kw, p = input_setup(inp)
    s = Sim(m, x, v, **kw)
    stat = {"calls": 0, "lists_equal": True, "acc": [], "pot": 0.0}

    def hook(jx, jv, xi, vi, h2, dtr, out, rows):
        racc, rjrk, rpot, rlst = O.regf(m, jx, jv, h2, dtr, xi, vi, s.lmax, s.nnbmax)
        stat["calls"] += 1
        stat["lists_equal"] &= rows_as_sets(rows) == rows_as_sets(rlst)
        stat["pot"] = max(stat["pot"], float((np.abs(out[6] - rpot) / rpot).max()))
        stat["acc"].append(vec_rel_err(out[0:3].T, racc))

    t_wall = time.time()
    try:
        s.start()
        e0 = s.energy()
        s.hook = hook
        while s.t < 0.25 and time.time() - t_wall < 90.0:
            s.block_step("rule")
        s.hook = None
        t_end = s.t
    finally:
        s.close()
    ea = np.concatenate(stat["acc"])
    print(f"1k.inp: N={m.size} NNBOPT={kw['nnbopt']} LMAX={kw['lmax']} to t={t_end:.4f}: "
          f"{stat['calls']} regular calls, acc p99.9 {np.percentile(ea, 99.9):.2e}, "
          f"pot max {stat['pot']:.2e}, E0 {e0:.6f}", flush=True)
    assert stat["calls"] > 0 and stat["lists_equal"]
    assert np.percentile(ea, 99.9) <= 1e-6 and stat["pot"] <= 1e-6
    assert abs(e0 + 0.25) < 0.02
