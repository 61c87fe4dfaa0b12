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


# ------------------------------------------------------------------ regular corrector
def _nb_lists(x, v, rs2, dtr, idx, exclude_self):
    """FP64 velocity-criterion neighbour sets min(r^2, |dx + dtr dv|^2) < rs2 of the
    particles idx (0-based), ascending 0-based indices."""
    out = []
    for k, i in enumerate(idx):
        dx = x - x[i]
        dv = v - v[i]
        r2 = (dx * dx).sum(1)
        p = dx + dtr[k] * dv
        rp2 = (p * p).sum(1)
        sel = np.minimum(r2, rp2) < rs2[k]
        if exclude_self:
            sel[i] = False
        out.append(np.flatnonzero(sel))
    return out


def pair_forces(x, v, m, i, js):
    """FP64 force and jerk of bodies js on body i (the pair terms of regcor_gpu.F:334-353)."""
    if len(js) == 0:
        return np.zeros(3), np.zeros(3)
    a = x[js] - x[i]
    dv = v[js] - v[i]
    r2 = (a * a).sum(1)
    dr3i = m[js] / (r2 * np.sqrt(r2))
    drdp = 3.0 * (a * dv).sum(1) / r2
    f = (a * dr3i[:, None]).sum(0)
    j = ((dv - a * drdp[:, None]) * dr3i[:, None]).sum(0)
    return f, j


def regcor_case(seed=7, n=800, ni=300, lmax=160, nnbopt=20, ncm=0, kz38=1, nb_flag=1,
                m_flag=0, tidal4=0.0, frac_void=0.04, frac_empty_old=0.05,
                frac_small_step=0.25, frac_few=0.04, frac_same=0.3, ifirst=1, kz37=0,
                nfast=0, ks_old=0.0):
    """A regular block for the corrector: state S (oracle layout), parameters P,
    1-based addresses, the regf output force [7][ni] (FP64 regular force and jerk over
    the non-neighbours) and the regf rows [ni][lmax] (0-based, self included).
    Particles n-ncm.. are c.m. bodies with binding energies h.  With ifirst > 1 the
    bodies 1..ifirst-1 stand for KS components: they are outside the j-set (regf rows
    are positions in IFIRST..NTOT), in no list and in no block.  Steps are powers of two
    commensurate with TIME = 1, so every branch of the step rule is reachable.
    kz37 = 1 with nfast > 0: LISTV holds nfast fast bodies (singles above the KS
    components, and some c.m. bodies), each pair has a KS separation r and CMSEP2 =
    1e4 (CHECKL).  ks_old: that fraction of the old lists also holds a few KS
    components J < IFIRST (needs IFIRST - 1 <= 2 ncm), which the new lists lose
    (FPCORR's c.m. branch)."""
    rng = np.random.default_rng(seed)
    m, x, v = ics.plummer(n, seed)
    nmax = n
    time = 1.0
    listv = None
    if nfast > 0:
        lo = max(ifirst, 2 * ncm + 1)                 # above every KS component
        pool = np.r_[np.arange(lo, n - ncm + 1), np.arange(n - ncm + 1, n + 1)[: ncm // 4]]
        fast = rng.choice(pool, min(nfast, pool.size), replace=False)
        v[fast - 1] *= 5.0
        listv = np.r_[len(fast), fast].astype(np.int32)
    dtmin = 2.0 ** -14
    f0 = ifirst - 1                                   # first j-set body, 0-based
    addr0 = np.sort(f0 + rng.choice(n - f0, ni, replace=False))
    kdt = rng.integers(3, 8, n)
    dtr_all = 2.0 ** (-kdt.astype(float))
    step = dtr_all / 2.0 ** rng.integers(0, 4, n)
    small = rng.random(n) < frac_small_step
    step[small] = dtmin                    # STEP < SMIN = 2 DTMIN
    rs = ics.rs_for_counts(x, np.arange(n), np.full(n, nnbopt))
    few = rng.random(n) < frac_few
    rs[few] = ics.rs_for_counts(x, np.flatnonzero(few), np.full(few.sum(), 3))
    nnbmax = min(n // 2, lmax - 50)
    # old lists at a perturbed radius (the previous regular step)
    rs_old = rs * rng.uniform(0.85, 1.15, n)
    x_prev = x - dtr_all[:, None] * v                 # positions one regular step back
    old = _nb_lists(x_prev, v, rs_old ** 2, dtr_all, np.arange(n), exclude_self=True)
    old = [o[o >= f0] for o in old]
    lst = np.zeros((nmax, lmax), np.int32)
    empty_old = rng.random(n) < frac_empty_old
    ks = (rng.random(n) < ks_old) if ks_old > 0 and ifirst > 1 else np.zeros(n, bool)
    for i in range(n):
        js = old[i][:nnbmax] if not empty_old[i] else old[i][:0]
        if ks[i] and len(js) > 0:
            comp = rng.choice(ifirst - 1, rng.integers(1, 4), replace=False)   # 0-based
            js = np.unique(np.r_[comp, js])[:nnbmax]
        lst[i, 0] = len(js)
        lst[i, 1:1 + len(js)] = js + 1
    # new rows in regf form: 0-based positions, self included; void rows hold self only
    dtr = dtr_all[addr0]
    new = _nb_lists(x, v, rs[addr0] ** 2, dtr, addr0, exclude_self=False)
    new = [o[o >= f0] for o in new]
    rows = np.zeros((ni, lmax), np.int32)
    void = rng.random(ni) < frac_void
    force = np.zeros((7, ni))
    for k, i in enumerate(addr0):
        js = np.array([i]) if void[k] else new[k][:nnbmax]
        rows[k, 0] = len(js)
        rows[k, 1:1 + len(js)] = js - f0             # regf rows: positions in the j-set
        non = np.setdiff1d(np.arange(n), js)
        f, j = pair_forces(x, v, m, i, non)
        force[0:3, k] = f
        force[3:6, k] = j
    # a fraction of the block keeps its list (no neighbour change)
    same = rng.random(ni) < frac_same
    for k, i in enumerate(addr0):
        if same[k] and not void[k]:
            js = rows[k, 1:1 + rows[k, 0]] + f0
            js = js[js != i]
            lst[i, :] = 0
            lst[i, 0] = len(js)
            lst[i, 1:1 + len(js)] = js + 1
    # state of the previous regular / irregular steps (arbitrary but smooth values)
    fi_old = np.zeros((3, n))
    fidot_old = np.zeros((3, n))
    for i in range(n):
        fi_old[:, i], fidot_old[:, i] = pair_forces(x, v, m, i, lst[i, 1:1 + lst[i, 0]] - 1)
    S = dict(
        x=x.T.copy(), xdot=v.T.copy(), body=m.copy(),
        x0=(x - 0.01 * v).T.copy(), x0dot=(v * 0.999).T.copy(),
        f=0.5 * rng.normal(size=(3, n)), fdot=rng.normal(size=(3, n)) / 6.0,
        t0=time - step, h=-rng.uniform(0.5, 5.0, max(ncm, 1)),
        step=step, timenw=time + step, fi=fi_old, fidot=fidot_old,
        fr=rng.normal(size=(3, n)) * 0.3, frdot=rng.normal(size=(3, n)) * 0.3,
        d2=rng.normal(size=(3, n)), d3=rng.normal(size=(3, n)),
        t0r=time - dtr_all, stepr=dtr_all.copy(), rs=rs.copy(), list=lst,
        xn=np.zeros((3, n)), xndot=np.zeros((3, n)), d0=np.zeros((3, n)),
        d0r=np.zeros((3, n)), d1r=np.zeros((3, n)), d2r=np.zeros((3, n)),
        d3r=np.zeros((3, n)))
    # regular force history of the block: from smooth (the force one step back along
    # its derivative) to unrelated, so that the new step can grow, stay or shrink
    eps = 10.0 ** rng.uniform(-6, 0, ni)
    for k, i in enumerate(addr0):
        fr = force[0:3, k] - dtr[k] * force[3:6, k]
        S["fr"][:, i] = fr + eps[k] * np.abs(fr).max() * rng.normal(size=3)
        S["frdot"][:, i] = force[3:6, k] * (1.0 + eps[k] * rng.normal(size=3))
    P = dict(time=time, dtmin=dtmin, smin=2.0 * dtmin, smax=0.125, etar=0.02, rscale=0.8,
             rc2=0.05, rsfac=12.0, znbmin=0.2 * nnbopt, bodym=1.0 / n, tidal4=tidal4,
             rdens=[0.01, -0.02, 0.005], n=n - ncm, ifirst=ifirst, nnbopt=nnbopt, nnbmax=nnbmax,
             nb_flag=nb_flag, m_flag=m_flag, kz38=kz38, kz37=kz37, cmsep2=1.0e4)
    if listv is not None:
        S["listv"] = listv
        S["r"] = 10.0 ** rng.uniform(-3.5, -1.5, max(ncm, 1))
    return S, P, addr0 + 1, force, rows


def copy_state(S):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in S.items()}


# ------------------------------------------------------------------ energy with KS pairs
def energy_case(seed=31, nsing=600, npairs=40, lmax=200, rs=0.25, rpert=0.5, full=False):
    """Bodies in common6.h order: KS components 1..2P, singles 2P+1..N, centres of mass
    N+1..N+P (masses m1 + m2, positions and velocities the mass-weighted means).  LIST
    rows: singles list singles and c.m. bodies within rs (ascending, so c.m. bodies come
    last); each pair's first component lists its perturbers (singles and other c.m.
    bodies) within rpert.  full=True lists every body instead.  Returns
    (x [nmax][3], xdot, body, lst, n, npairs, ht)."""
    rng = np.random.default_rng(seed)
    nb = nsing + npairs
    m, x, v = ics.plummer(nb, seed)
    P = npairs
    n = 2 * P + nsing
    nmax = n + P
    X = np.zeros((nmax, 3))
    V = np.zeros((nmax, 3))
    M = np.zeros(nmax)
    X[2 * P:n], V[2 * P:n], M[2 * P:n] = x[:nsing], v[:nsing], m[:nsing]
    X[n:], V[n:], M[n:] = x[nsing:], v[nsing:], m[nsing:]
    for p in range(P):
        q = rng.uniform(0.2, 0.8)
        mc = M[n + p]
        m1, m2 = q * mc, (1 - q) * mc
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        a = 10 ** rng.uniform(-3.5, -2)
        w = rng.normal(size=3) * 0.3
        X[2 * p] = X[n + p] + (m2 / mc) * a * u
        X[2 * p + 1] = X[n + p] - (m1 / mc) * a * u
        V[2 * p] = V[n + p] + (m2 / mc) * w
        V[2 * p + 1] = V[n + p] - (m1 / mc) * w
        M[2 * p], M[2 * p + 1] = m1, m2
    lst = np.zeros((nmax, lmax), np.int32)
    bodies = np.arange(2 * P + 1, nmax + 1)          # singles and c.m. bodies, 1-based
    for I in range(2 * P + 1, n + 1):
        d = X[bodies - 1] - X[I - 1]
        sel = bodies[((d * d).sum(1) < rs * rs) | full]
        sel = sel[sel != I][:lmax - 1]
        lst[I - 1, 0] = len(sel)
        lst[I - 1, 1:1 + len(sel)] = sel
    for p in range(P):
        I = n + p + 1
        d = X[bodies - 1] - X[I - 1]
        sel = bodies[((d * d).sum(1) < rpert * rpert) | full]
        sel = sel[sel != I][:lmax - 1]
        lst[2 * p, 0] = len(sel)
        lst[2 * p, 1:1 + len(sel)] = sel
    ht = -rng.uniform(1.0, 50.0, P)
    return X, V, M, lst, n, P, ht
