"""Pins the CPU restatement of the regular corrector (oracle/regcor.py,
regcor_gpu.F + fpcorr.F) with checks that do not retype its formulas: set
differences, conservation of the total force when neighbours change sides, the
Hermite corrector on an exactly cubic force, the second and third force derivatives
of fpcorr.F against an independent power-series expansion, the REAL*4 literals of
the RS rule, the SMIN retention and the step rules."""
import math

import numpy as np
import pytest

import regcor as R
from helpers import copy_state, pair_forces, regcor_case

f32 = np.float32


def test_merge_lists_is_set_difference():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = sorted(rng.choice(60, rng.integers(0, 30), replace=False).tolist())
        b = sorted(rng.choice(60, rng.integers(0, 30), replace=False).tolist())
        lo, ga = R.merge_lists(a, b)
        assert lo == sorted(set(a) - set(b))
        assert ga == sorted(set(b) - set(a))


@pytest.mark.parametrize("small", [0.0, 0.3])
def test_neighbour_changes_keep_the_total_force(small):
    """FI (old list) + the moved forces = the force of the final list; FR = FREG of the
    complement; so FI + FR is the full direct sum, for losses, gains and retained
    small-step neighbours alike (regcor_gpu.F:298-415)."""
    S, P, addr, force, rows = regcor_case(seed=3, frac_small_step=small, frac_void=0.0, kz38=0)
    m = S["body"]
    x = S["x"].T
    v = S["xdot"].T
    S1 = copy_state(S)
    C = R.regcor_block(S1, P, addr, force, rows)
    if small:
        assert C["nbsmin"] > 0
    assert C["nbflux"] > 0
    n = len(m)
    for I in addr:
        i = I - 1
        row = S1["list"][i]
        final = row[1:row[0] + 1] - 1
        fi, fj = pair_forces(x, v, m, i, final)
        tot_f, tot_j = pair_forces(x, v, m, i, np.setdiff1d(np.arange(n), [i]))
        sc = np.abs(tot_f).max() + 1e-300
        sj = np.abs(tot_j).max() + 1e-300
        assert np.allclose(S1["fi"][:, i], fi, rtol=0, atol=1e-11 * sc)
        assert np.allclose(S1["fidot"][:, i], fj, rtol=0, atol=1e-11 * sj)
        assert np.allclose(S1["fi"][:, i] + S1["fr"][:, i], tot_f, rtol=0, atol=1e-11 * sc)
        assert np.allclose(S1["fidot"][:, i] + S1["frdot"][:, i], tot_j, rtol=0, atol=1e-11 * sj)


def test_no_neighbours_central_mass():
    """NNB = 0: the old neighbours leave and a mass 0.01 BODYM (REAL*4 0.01) at the
    origin is added to FI; RS grows by the REAL*4 factor 1.1 when moving inwards
    (regcor_gpu.F:95-127)."""
    S, P, addr, force, rows = regcor_case(seed=5, frac_void=0.3)
    S1 = copy_state(S)
    C = R.regcor_block(S1, P, addr, force, rows)
    assert C["nbvoid"] > 0
    x = S["x"].T
    v = S["xdot"].T
    checked_in = checked_out = 0
    for k, I in enumerate(addr):
        i = I - 1
        if rows[k][0] != 0:
            continue
        r2 = x[i] @ x[i]
        fc = -float(f32(0.01)) * P["bodym"] / (r2 * math.sqrt(r2)) * x[i]
        sc = np.abs(S["fi"][:, i]).max() + np.abs(fc).max()
        assert np.allclose(S1["fi"][:, i], fc, rtol=0, atol=1e-12 * sc)
        if x[i] @ v[i] < 0:
            assert S1["rs"][i] == float(f32(1.1)) * S["rs"][i]
            checked_in += 1
        else:
            assert S1["rs"][i] == max(0.75 * S["rs"][i], float(f32(0.1)) * P["rscale"])
            checked_out += 1
    assert checked_in and checked_out


def _single(time=1.0, dtr=0.0625, nlist=(2, 3), old=(2, 3), nmax=4, lmax=16):
    """Particle 1 with fixed neighbours; the others far away; all arrays zero."""
    z3 = np.zeros((3, nmax))
    x = np.zeros((3, nmax))
    x[0, :] = 30.0 + np.arange(nmax)
    x[:, :4] = [[0.0, 0.05, -0.04, 30.0], [0.0, 0.02, 0.03, 0.0], [0.0, 0.0, 0.01, 0.0]]
    xdot = np.zeros((3, nmax))
    xdot[:, :4] = [[0.1, -0.2, 0.1, 0.0], [0.0, 0.1, 0.0, 0.0], [0.05, 0.0, 0.2, 0.0]]
    body = np.full(nmax, 1e-3)
    body[:3] = [1e-3, 2e-3, 1.5e-3]
    S = dict(x=x, xdot=xdot, body=body,
             x0=z3.copy(), x0dot=z3.copy(), f=z3.copy(), fdot=z3.copy(), t0=np.full(nmax, time),
             h=np.array([-1.0]), step=np.full(nmax, dtr), timenw=np.zeros(nmax),
             fi=z3.copy(), fidot=z3.copy(), fr=z3.copy(), frdot=z3.copy(), d2=z3.copy(),
             d3=z3.copy(), t0r=np.full(nmax, time - dtr), stepr=np.full(nmax, dtr),
             rs=np.full(nmax, 0.1), list=np.zeros((nmax, lmax), np.int32), xn=z3.copy(),
             xndot=z3.copy(), d0=z3.copy(), d0r=z3.copy(), d1r=z3.copy(), d2r=z3.copy(),
             d3r=z3.copy())
    S["list"][0, 0] = len(old)
    S["list"][0, 1:1 + len(old)] = old
    row = np.zeros((1, lmax), np.int32)
    js = sorted([0] + [j - 1 for j in nlist])          # regf form: 0-based, self included
    row[0, 0] = len(js)
    row[0, 1:1 + len(js)] = js
    P = dict(time=time, dtmin=2.0 ** -14, smin=2.0 ** -13, smax=0.125, etar=0.02, rscale=1.0,
             rc2=1e-4, rsfac=10.0, znbmin=1.0, bodym=1e-3, tidal4=0.0, rdens=[0.0, 0.0, 0.0],
             n=nmax, ifirst=1, nnbopt=2, nnbmax=8, nb_flag=1, m_flag=0, kz38=0)
    return S, P, row


def test_hermite_corrector_on_a_cubic_force():
    """Regular force F(t) = a0 + a1 s + a2 s^2/2 + a3 s^3/6 (s from T0R): the corrector
    must return D3R = F''' and D2R = F''(TIME), and move X0, X0DOT by the double and
    single integrals of a2 s^2/2 + a3 s^3/6 over the step (regcor_gpu.F:417-530).
    XN carries the REAL*4 factor 0.6 (relative 4e-8 on its first term)."""
    rng = np.random.default_rng(11)
    dtr = 0.0625
    S, P, row = _single(dtr=dtr)
    a0, a1, a2, a3 = rng.normal(size=(4, 3))
    F = lambda s: a0 + a1 * s + a2 * s * s / 2 + a3 * s ** 3 / 6
    Fd = lambda s: a1 + a2 * s + a3 * s * s / 2
    S["fr"][:, 0] = F(0.0)
    S["frdot"][:, 0] = Fd(0.0)
    S["x0"][:, 0] = [0.3, -0.1, 0.2]
    S["x0dot"][:, 0] = [0.01, 0.02, -0.03]
    force = np.zeros((7, 1))
    force[0:3, 0] = F(dtr)
    force[3:6, 0] = Fd(dtr)
    R.regcor_block(S, P, [1], force, row)
    assert np.allclose(S["d3r"][:, 0], a3, rtol=1e-9, atol=1e-9)
    assert np.allclose(S["d2r"][:, 0], a2 + a3 * dtr, rtol=1e-9, atol=1e-9)
    dv = a2 * dtr ** 3 / 6 + a3 * dtr ** 4 / 24
    dx = a2 * dtr ** 4 / 24 + a3 * dtr ** 5 / 120
    assert np.allclose(S["xndot"][:, 0] - [0.01, 0.02, -0.03], dv, rtol=1e-9, atol=1e-15)
    assert np.allclose(S["xn"][:, 0] - [0.3, -0.1, 0.2], dx, rtol=1e-6, atol=1e-15)
    assert S["t0r"][0] == 1.0
    assert np.array_equal(S["fr"][:, 0], force[0:3, 0])
    assert np.array_equal(S["d1r"][:, 0], force[3:6, 0])


def _series_force_derivs(r, v, a, j, m):
    """2nd and 3rd time derivatives at t = 0 of m r(t)/|r(t)|^3 with
    r(t) = r + v t + a t^2/2 + j t^3/6, by truncated power-series arithmetic."""
    c = [np.asarray(r, float), np.asarray(v, float), np.asarray(a, float) / 2,
         np.asarray(j, float) / 6]
    s = [sum(c[p] @ c[k - p] for p in range(k + 1)) for k in range(4)]     # |r|^2
    u = [0.0] + [s[k] / s[0] for k in range(1, 4)]                        # s = s0 (1 + u)

    def mul(x, y):
        return [sum(x[p] * y[k - p] for p in range(k + 1)) for k in range(4)]

    g = [1.0, 0.0, 0.0, 0.0]
    term = [1.0, 0.0, 0.0, 0.0]
    coef = 1.0
    for n in range(1, 4):                                                  # (1+u)^(-3/2)
        coef *= (-1.5 - (n - 1)) / n
        term = mul(term, u)
        g = [g[k] + coef * term[k] for k in range(4)]
    g = [gk * s[0] ** -1.5 for gk in g]
    f = [m * sum(c[p] * g[k - p] for p in range(k + 1)) for k in range(4)]
    return 2.0 * f[2], 6.0 * f[3]


def test_fpcorr_derivatives_match_power_series():
    """fpcorr.F:36-74 gives the second and third derivative of the pair force along
    the relative motion with acceleration A(7:9) and jerk A(10:12)."""
    rng = np.random.default_rng(4)
    S, P, row = _single()
    P["kz38"] = 1
    j = 2                                      # body 3 (0-based 2)
    S["t0"][j] = 0.9
    S["f"][:, j] = rng.normal(size=3) * 0.3
    S["fdot"][:, j] = rng.normal(size=3) * 0.2
    S["f"][:, 0] = rng.normal(size=3) * 0.3
    S["fdot"][:, 0] = rng.normal(size=3) * 0.2
    S["x0dot"][:, j] = rng.normal(size=3) * 0.1
    xi = S["x"][:, 0]
    vi = S["xdot"][:, 0]
    s = P["time"] - S["t0"][j]
    r = S["x"][:, j] - xi
    v = (S["fdot"][:, j] * 3 * s + 2 * S["f"][:, j]) * s + S["x0dot"][:, j] - vi
    a = 2 * (S["fdot"][:, j] * 3 * s + S["f"][:, j] - S["f"][:, 0])
    jj = 6 * (S["fdot"][:, j] - S["fdot"][:, 0])
    want2, want3 = _series_force_derivs(r, v, a, jj, S["body"][j])
    C = dict(ncmder=0)
    s2, s3 = R.fpcorr_terms(S, P, 0, list(xi), list(vi), [j + 1], 0, C)
    assert np.allclose(s2, want2, rtol=1e-10, atol=0)
    assert np.allclose(s3, want3, rtol=1e-10, atol=0)
    l2, l3 = R.fpcorr_terms(S, P, 0, list(xi), list(vi), [j + 1], 1, C)   # as a loss
    assert np.allclose(l2, -np.asarray(want2)) and np.allclose(l3, -np.asarray(want3))


def test_fpcorr_ks_component_uses_its_cm_derivatives():
    """A KS component J < IFIRST (bodies 1, 2 of pair 1; c.m. body N + KVEC(J) = 5):
    the terms are the 2nd and 3rd derivatives of its pair force along the motion with
    its own position and velocity and the c.m. body's acceleration 2 F and jerk
    6 FDOT (fpcorr.F:46-61), without the prediction of the single-body branch."""
    rng = np.random.default_rng(6)
    S, P, row = _single(nmax=6)
    P.update(ifirst=3, n=4)
    i = 2                                      # particle 3 (a single)
    S["x"][:, :3] = [[0.05, 0.0502, 0.0], [0.02, 0.0199, 0.0], [0.0, 0.0001, 0.0]]
    S["t0"][:] = 0.5                           # must not enter the component branch
    for b in (0, 1, 4, i):
        S["f"][:, b] = rng.normal(size=3) * 0.3
        S["fdot"][:, b] = rng.normal(size=3) * 0.2
    xi, vi = S["x"][:, i], S["xdot"][:, i]
    for J in (1, 2):
        jj = J - 1
        r = S["x"][:, jj] - xi
        v = S["xdot"][:, jj] - vi
        a = 2 * (S["f"][:, 4] - S["f"][:, i])
        jk = 6 * (S["fdot"][:, 4] - S["fdot"][:, i])
        want2, want3 = _series_force_derivs(r, v, a, jk, S["body"][jj])
        C = dict(ncmder=0)
        s2, s3 = R.fpcorr_terms(S, P, i, list(xi), list(vi), [J], 0, C)
        assert C["ncmder"] == 1
        assert np.allclose(s2, want2, rtol=1e-10, atol=0)
        assert np.allclose(s3, want3, rtol=1e-10, atol=0)


def _fast_case(J_pos, J_vel, listv, nmax=8):
    """Particle 1 with neighbours 2, 3 and RS = 0.1; bodies in listv placed at J_pos
    (relative to particle 1) with velocities J_vel."""
    S, P, row = _single(nlist=(2, 3), old=(2, 3), nmax=nmax)
    for J, (dx, dv) in zip(listv, zip(J_pos, J_vel)):
        S["x"][:, J - 1] = S["x"][:, 0] + np.asarray(dx)
        S["xdot"][:, J - 1] = S["xdot"][:, 0] + np.asarray(dv)
    S["listv"] = np.r_[len(listv), listv].astype(np.int32)
    S["r"] = np.array([1e-3])
    P.update(kz37=1, cmsep2=1.0e4)
    return S, P, row


def test_checkl_adds_the_first_approaching_intruder():
    """KZ(37) = 1 (checkl.F:243-340): LISTV = [6 receding, 7 outside 2 RS, 8 impact
    parameter above RS, 2 already a neighbour, 5 approaching, 4 approaching]: only 5
    joins the list (one per call), its force leaves FR and reaches FI through the gain,
    so FI is the direct sum over the final list and FI + FR is unchanged; NBFAST = 1.
    With KZ(37) = 0 the list stays as it was."""
    pos = [[0.15, 0, 0], [0.25, 0, 0], [0.15, 0.12, 0], [0.05, 0.02, 0], [0.12, 0.0, 0.03],
           [0.1, -0.1, 0]]
    vel = [[1.0, 0, 0], [-1.0, 0, 0], [-1.0, 0, 0], [0.1, 0.1, 0.0], [-2.0, 0.1, 0.0],
           [-1.0, 1.0, 0.0]]
    lv = [6, 7, 8, 2, 5, 4]
    for kz37 in (0, 1):
        S, P, row = _fast_case(pos, vel, lv)
        P["kz37"] = kz37
        if kz37 == 1:
            S["x"][:, 1] = S["x"][:, 0] + np.array(pos[3])
            S["xdot"][:, 1] = S["xdot"][:, 0] + np.array(vel[3])
        x, v, m = S["x"].T.copy(), S["xdot"].T.copy(), S["body"]
        force = np.zeros((7, 1))
        force[0:3, 0], force[3:6, 0] = pair_forces(x, v, m, 0, np.array([3, 4, 5, 6, 7]))
        fi0 = S["fi"][:, 0].copy()
        C = R.regcor_block(S, P, [1], force, row)
        final = S["list"][0, 1:S["list"][0, 0] + 1] if C["nbflux"] else np.array([2, 3])
        if kz37 == 0:
            assert C["nbfast"] == 0 and list(row[0, :3]) == [2, 2, 3]
            continue
        assert C["nbfast"] == 1
        assert list(row[0, :4]) == [3, 2, 3, 5]
        assert list(final) == [2, 3, 5]
        f5, j5 = pair_forces(x, v, m, 0, np.array([4]))
        sf, sj = np.abs(force[0:3, 0]).max(), np.abs(force[3:6, 0]).max()
        assert np.allclose(S["fr"][:, 0], force[0:3, 0] - f5, rtol=0, atol=1e-14 * sf)
        assert np.allclose(S["frdot"][:, 0], force[3:6, 0] - j5, rtol=0, atol=1e-13 * sj)
        assert np.allclose(S["fi"][:, 0], fi0 + f5, rtol=0, atol=1e-14 * sf)


def test_checkl_cm_particle_takes_component_terms():
    """A c.m. particle I > N near an intruder J (single, inside CMSEP2 R^2): the force
    leaving FREG is the mass-weighted sum over I's two KS components
    m_J m_L / m_I (x_J - x_L) / r^3 (checkl.F:282-336), not the point-mass term."""
    nmax = 8
    S, P, row = _single(nlist=(6,), old=(6,), nmax=nmax)
    # bodies 1, 2: components of pair 1 (c.m. = body 7 = N + 1 with N = 6), I = 7
    P.update(n=6, ifirst=3, kz37=1, cmsep2=1.0e4)
    m1, m2 = 1.2e-3, 0.8e-3
    S["body"][[0, 1, 6]] = [m1, m2, m1 + m2]
    cm = np.array([0.2, 0.1, -0.05])
    d = np.array([2e-3, -1e-3, 5e-4])
    S["x"][:, 6] = cm
    S["x"][:, 0] = cm + m2 / (m1 + m2) * d
    S["x"][:, 1] = cm - m1 / (m1 + m2) * d
    S["xdot"][:, 6] = [0.1, 0.0, 0.0]
    S["xdot"][:, 0] = S["xdot"][:, 6] + [0.0, 1.0, 0.0]
    S["xdot"][:, 1] = S["xdot"][:, 6] - m1 / m2 * np.array([0.0, 1.0, 0.0])
    S["x"][:, 5] = cm + [0.05, 0.0, 0.0]            # the old and new neighbour 6
    S["x"][:, 4] = cm + [0.08, 0.01, 0.0]           # intruder 5, approaching
    S["xdot"][:, 4] = S["xdot"][:, 6] + [-1.0, 0.0, 0.0]
    S["listv"] = np.array([1, 5], np.int32)
    S["r"] = np.array([np.linalg.norm(d)])          # CMSEP2 R^2 = 0.05 > 0.08^2
    S["rs"][6] = 0.1
    S["h"] = np.array([-5.0])
    S["list"][:, :] = 0
    S["list"][6, 0:2] = [1, 6]
    row = np.zeros((1, S["list"].shape[1]), np.int32)
    row[0, 0:3] = [2, 3, 4]             # regf form: positions from IFIRST = 3: bodies 6, 7 (self)
    force = np.zeros((7, 1))
    force[0:3, 0] = [0.3, -0.2, 0.1]
    force[3:6, 0] = [0.01, 0.02, -0.03]
    C = R.regcor_block(S, P, [7], force, row)
    assert C["nbfast"] == 1
    assert list(row[0, :3]) == [2, 5, 6]
    want_f = np.zeros(3)
    want_j = np.zeros(3)
    for L, mL in ((0, m1), (1, m2)):
        a = S["x"][:, 4] - S["x"][:, L]
        dv = S["xdot"][:, 4] - S["xdot"][:, L]
        r2 = a @ a
        w = S["body"][4] * mL / (m1 + m2) / (r2 * np.sqrt(r2))
        want_f += a * w
        want_j += (dv - 3 * (a @ dv) / r2 * a) * w
    assert np.allclose(S["fr"][:, 6], force[0:3, 0] - want_f, rtol=0, atol=1e-13 * np.abs(want_f).max())
    assert np.allclose(S["frdot"][:, 6], force[3:6, 0] - want_j, rtol=0,
                       atol=1e-12 * np.abs(want_j).max())
    pf, pj = pair_forces(S["x"].T, S["xdot"].T, S["body"], 6, np.array([4]))
    assert not np.allclose(pf, want_f, rtol=1e-6)      # the point mass would differ


def test_fpcorr_moves_derivatives_between_parts():
    """FPCORR adds to D2, D3 what it takes from D2R, D3R (fpcorr.F:93-97): the sums
    equal those of the run without it, and KZ(38) = 3 counts losses only."""
    base = regcor_case(seed=9, kz38=0)
    outs = {}
    for kz in (0, 1, 2, 3):
        S, P, addr, force, rows = base
        S1 = copy_state(S)
        P1 = dict(P, kz38=kz)
        C = R.regcor_block(S1, P1, addr, force, rows.copy())
        outs[kz] = (S1, C)
    S0 = outs[0][0]
    ia = base[2] - 1
    for kz in (1, 2, 3):
        S1, C = outs[kz]
        assert C["nbcorr"] > 0
        for a, b in (("d2", "d2r"), ("d3", "d3r")):
            tot0 = S0[a][:, ia] + S0[b][:, ia]
            tot1 = S1[a][:, ia] + S1[b][:, ia]
            assert np.allclose(tot1, tot0, rtol=1e-12, atol=1e-12 * np.abs(tot0).max())
    assert outs[0][1]["nbcorr"] == 0
    assert outs[3][1]["nbflux"] < outs[1][1]["nbflux"]


def test_rs_rule_uses_single_precision_literals():
    """NNB = NNB0 = 4 < ZNBMIN = 5 at RI = 10 RSCALE with NNBOPT = 100, NBAVE = 250:
    FAC = 3.4, A3 = 13.6 gives NBP = 13 > ZNBMIN, the outer-shell factor 0.1 makes
    A4 = 0.34, clamped to 1 - A5 = 0.75, so A1 < 1 is replaced by the REAL*4 literal
    1.05 (regcor_gpu.F:198-200): RS grows by exactly float32(1.05)."""
    S, P, row = _single(nlist=(2, 3, 4, 5), old=(2, 3, 4, 5), nmax=6)
    S["x"][:, 0] = [10.0, 0.0, 0.0]
    for k, d in enumerate(([0.05, 0, 0], [0, 0.05, 0], [0, 0, 0.05], [-0.05, 0, 0])):
        S["x"][:, 1 + k] = S["x"][:, 0] + d
    S["xdot"][:, :] = 0.0
    S["xdot"][1, 0] = 0.1                      # tangential: RIDOT = 0
    P.update(nnbopt=100, znbmin=5.0, nnbmax=300, rscale=1.0, rsfac=10.0)
    rs0 = S["rs"][0]
    force = np.zeros((7, 1))
    force[0:6, 0] = [-0.01, 0.002, 0.0, 0.001, -0.003, 0.0005]
    C = R.regcor_block(S, P, [1], force, row)
    assert C["nbflux"] == 0 and C["nbvoid"] == 0
    assert S["rs"][0] == float(f32(1.05)) * rs0
    assert S["rs"][0] != 1.05 * rs0


def test_small_step_neighbour_is_retained():
    """Old list {2, 3}, new list {3}: body 2 with STEP < SMIN inside 2 RS stays a
    neighbour, its force leaves FREG, and the list is not rewritten; with a larger
    step it is lost (regcor_gpu.F:298-350)."""
    for small in (True, False):
        S, P, row = _single(nlist=(3,), old=(2, 3))
        S["step"][1] = 2.0 ** -14 if small else 2.0 ** -10
        force = np.zeros((7, 1))
        force[0:3, 0] = [0.1, 0.2, 0.3]
        force[3:6, 0] = [0.01, 0.02, 0.03]
        C = R.regcor_block(S, P, [1], force, row)
        x = S["x"].T
        v = S["xdot"].T
        f2, j2 = pair_forces(x, v, S["body"], 0, np.array([1]))
        if small:
            assert C["nbsmin"] == 1 and C["nbflux"] == 0
            assert list(S["list"][0, :3]) == [2, 2, 3]
            assert list(row[0, :3]) == [2, 2, 3]
            assert np.allclose(S["fr"][:, 0], force[0:3, 0] - f2, rtol=1e-14)
            assert np.allclose(S["frdot"][:, 0], force[3:6, 0] - j2, rtol=1e-14)
        else:
            assert C["nbsmin"] == 0 and C["nbflux"] == 1
            assert list(S["list"][0, :2]) == [1, 3]
            assert np.allclose(S["fi"][:, 0], -f2, rtol=1e-14)


def test_step_rules():
    """New regular steps (regcor_gpu.F:624-740) stay commensurate with TIME, within a
    factor 4 below and SMAX above, take every branch of the rule on this block, and an
    empty old list divides STEP by 8 with TIMENW = T0 + STEP."""
    S, P, addr, force, rows = regcor_case(seed=21)
    S1 = copy_state(S)
    R.regcor_block(S1, P, addr, force, rows)
    ia = addr - 1
    old, new = S["stepr"][ia], S1["stepr"][ia]
    assert np.all(np.fmod(P["time"], new) == 0.0)
    assert np.all(new <= P["smax"]) and np.all(new >= old / 4)
    assert np.any(new > old) and np.any(new < old) and np.any(new == old)
    lg = np.log2(new)
    assert np.all(lg == np.round(lg))
    empty = S["list"][ia, 0] == 0
    assert empty.any()
    assert np.array_equal(S1["step"][ia[empty]], S["step"][ia[empty]] / 8)
    assert np.array_equal(S1["timenw"][ia[empty]], S["t0"][ia[empty]] + S1["step"][ia[empty]])
    assert np.array_equal(S1["step"][ia[~empty]], S["step"][ia[~empty]])
    assert np.all(S1["t0r"][ia] == P["time"])
