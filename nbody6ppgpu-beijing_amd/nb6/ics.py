"""Seeded synthetic star-cluster initial conditions in N-body units (G = M = 1,
E = -1/4), for tests and benchmarks of the regular-force path.

The reference ships only a Plummer generator inside its Fortran set-up
(setup.F:59-100, cut at r > 10 at setup.F:64) and reads everything else from
dat.10 (data.F:36-39); it has no King generator.  These generators are written
for this library:

* ``plummer``   -- Plummer sphere, positions by inverse mass fraction, speeds by
                   von Neumann rejection on q^2 (1-q^2)^(7/2) (Aarseth, Henon &
                   Wielen 1974), radii cut at r > 10 like setup.F:64.
* ``king``      -- King (1966) model of central potential W0, from the ODE for
                   the dimensionless potential, velocities by rejection on the
                   lowered-Maxwellian distribution.
* ``kroupa_masses`` -- Kroupa (2001) two-segment IMF (alpha 1.3 / 2.3, break at
                   0.5 Msun) between 0.08 and 100 Msun, normalised to total 1.
* ``add_binaries`` -- splits a fraction of the systems into binaries with
                   log-uniform semi-major axes in AU and thermal eccentricities.

All randomness comes from ``numpy.random.Generator(PCG64(seed))``.
"""
from __future__ import annotations

import math

import numpy as np

AU_PER_PC = 206264.806


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _random_unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    s = np.sqrt(1.0 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)


def kroupa_masses(n: int, seed: int, mlo: float = 0.08, mhi: float = 100.0,
                  mbreak: float = 0.5) -> np.ndarray:
    """Kroupa (2001) IMF: dN/dm ~ m^-1.3 below ``mbreak``, m^-2.3 above; total 1."""
    rng = _rng(seed)
    a1, a2 = 1.3, 2.3
    # segment integrals of the (continuous) broken power law
    def seg(a, lo, hi):
        return (hi ** (1 - a) - lo ** (1 - a)) / (1 - a)
    c2 = mbreak ** (a2 - a1)            # continuity at the break
    w1 = seg(a1, mlo, mbreak)
    w2 = c2 * seg(a2, mbreak, mhi)
    u = rng.uniform(0.0, 1.0, n)
    pick1 = rng.uniform(0.0, 1.0, n) < w1 / (w1 + w2)
    m = np.empty(n)
    def inv(a, lo, hi, uu):
        return (lo ** (1 - a) + uu * (hi ** (1 - a) - lo ** (1 - a))) ** (1.0 / (1 - a))
    m[pick1] = inv(a1, mlo, mbreak, u[pick1])
    m[~pick1] = inv(a2, mbreak, mhi, u[~pick1])
    return m / m.sum()


def _to_nbody_units(m, x, v, pot_energy):
    """Centre on the centre of mass and scale to G=M=1, E=-1/4 (virial Q=0.5)."""
    mt = m.sum()
    m = m / mt
    x = x - (m[:, None] * x).sum(0)
    v = v - (m[:, None] * v).sum(0)
    ke = 0.5 * (m * (v * v).sum(1)).sum()
    # pot_energy is given for total mass mt; rescale to total mass 1
    pe = pot_energy / (mt * mt)
    # virial equilibrium: scale velocities so that 2T = -W, then lengths so E = -1/4
    qv = math.sqrt(-0.5 * pe / ke)
    v = v * qv
    e = 0.5 * pe                            # T = -W/2  =>  E = W/2
    s = e / -0.25                           # lengths scale by s, energies by 1/s
    x = x * s
    v = v / math.sqrt(s)
    return m, x, v


def plummer(n: int, seed: int, masses: np.ndarray | None = None, rcut: float = 10.0):
    """Plummer sphere in N-body units.  Returns (m[n], x[n,3], v[n,3]) as float64."""
    rng = _rng(seed)
    m = np.full(n, 1.0 / n) if masses is None else np.asarray(masses, float).copy()
    # radii in units of the Plummer scale a=1 via inverse cumulative mass
    r = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        xm = rng.uniform(0.0, 1.0, todo.size)
        rr = 1.0 / np.sqrt(np.maximum(xm, 1e-300) ** (-2.0 / 3.0) - 1.0)
        ok = rr <= rcut
        r[todo[ok]] = rr[ok]
        todo = todo[~ok]
    x = _random_unit_vectors(rng, n) * r[:, None]
    # speeds: q = v/vesc with g(q) = q^2 (1-q^2)^3.5 by rejection (max g < 0.1)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x1 = rng.uniform(0.0, 1.0, todo.size)
        x2 = rng.uniform(0.0, 0.1, todo.size)
        ok = x2 < x1 * x1 * (1.0 - x1 * x1) ** 3.5
        q[todo[ok]] = x1[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0) * (1.0 + r * r) ** -0.25
    v = _random_unit_vectors(rng, n) * (q * vesc)[:, None]
    # Plummer potential energy for total mass M, scale a=1:  W = -3 pi M^2 / 32
    mt = m.sum()
    pe = -3.0 * math.pi * mt * mt / 32.0
    return _to_nbody_units(m, x, v, pe)


def _king_profile(w0: float, nstep: int = 20000):
    """Integrate the King (1966) dimensionless potential W(r) (r in core radii,
    rho(W)/rho0 normalised) out to the tidal radius W = 0.
    Returns arrays r, W, M(r) (arbitrary units, monotone) and the potential energy
    in the same units."""
    def rho(w):
        w = np.maximum(w, 0.0)
        sw = np.sqrt(w)
        return np.exp(w) * _erf(sw) - np.sqrt(4.0 * w / math.pi) * (1.0 + 2.0 * w / 3.0)

    rho0 = float(rho(np.array([w0]))[0])

    def deriv(r, y):
        w, dw = y
        if r == 0.0:
            return np.array([dw, -3.0 * float(rho(np.array([w]))[0]) / rho0])
        return np.array([dw, -9.0 * float(rho(np.array([w]))[0]) / rho0 - 2.0 * dw / r])

    # RK4 with step h; near r=0 use the series W = W0 - 1.5 r^2
    h = 2e-3
    rs, ws, dws = [0.0], [w0], [0.0]
    r, y = 0.0, np.array([w0, 0.0])
    while y[0] > 0.0 and r < 1e4:
        k1 = deriv(r, y)
        k2 = deriv(r + h / 2, y + h / 2 * k1)
        k3 = deriv(r + h / 2, y + h / 2 * k2)
        k4 = deriv(r + h, y + h * k3)
        ynew = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        if ynew[0] <= 0.0:
            # linear interpolation to W = 0
            f = y[0] / (y[0] - ynew[0])
            r = r + f * h
            y = y + f * (ynew - y)
            y[0] = 0.0
            rs.append(r); ws.append(0.0); dws.append(y[1])
            break
        r += h
        y = ynew
        rs.append(r); ws.append(y[0]); dws.append(y[1])
        if r > 2.0:
            h = min(0.02, h * 1.001)
    r = np.array(rs); w = np.array(ws); dw = np.array(dws)
    mass = -r * r * dw                      # M(r) up to a constant (G, 4 pi rho0 r_c^3/9)
    # potential energy: W = -(1/2) int rho(phi) phi dV with phi = phi_t - W (sigma=1 units
    # where 9/(4 pi G rho0 rc^2) = 1/sigma^2 -> rho0 = 9/(4 pi)); phi_t = -M_t / r_t
    rho_r = rho(w) / rho0 * (9.0 / (4.0 * math.pi))
    m_tot = mass[-1]
    phi = -m_tot / r[-1] - w
    integrand = rho_r * phi * 4.0 * math.pi * r * r
    pe = 0.5 * float(np.trapezoid(integrand, r))
    return r, w, mass, m_tot, pe


def _erf(x):
    from scipy.special import erf
    return erf(x)


def king(n: int, seed: int, w0: float, masses: np.ndarray | None = None):
    """King (1966) model with central potential ``w0``, N-body units."""
    rng = _rng(seed)
    r_tab, w_tab, m_tab, m_tot, pe = _king_profile(w0)
    frac = m_tab / m_tot
    u = rng.uniform(0.0, 1.0, n)
    r = np.interp(u, frac, r_tab)
    w = np.interp(r, r_tab, w_tab)
    x = _random_unit_vectors(rng, n) * r[:, None]
    # speeds: f(v) ~ v^2 (exp(W - v^2/2) - 1), 0 <= v <= sqrt(2W)  (sigma = 1)
    sp = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        wt = w[todo]
        vmax = np.sqrt(2.0 * wt)
        vv = rng.uniform(0.0, 1.0, todo.size) * vmax
        g = vv * vv * (np.exp(wt - 0.5 * vv * vv) - 1.0)
        # bound of g on [0, vmax]: v^2 e^W <= 2W e^W
        gmax = 2.0 * wt * np.exp(wt) + 1e-300
        ok = rng.uniform(0.0, 1.0, todo.size) * gmax < g
        sp[todo[ok]] = vv[ok]
        todo = todo[~ok]
    v = _random_unit_vectors(rng, n) * sp[:, None]
    m = np.full(n, 1.0 / n) if masses is None else np.asarray(masses, float).copy()
    # model units: total mass m_tot (with rho0 = 9/4pi, rc = sigma = G = 1); the
    # model potential energy scales with (sum m)^2 / m_tot^2 for the sampled masses
    mt = m.sum()
    pe_scaled = pe * (mt / m_tot) ** 2
    # velocities are in units of sigma for a model of mass m_tot; rescale to mt
    v = v * math.sqrt(mt / m_tot)
    return _to_nbody_units(m, x, v, pe_scaled)


def add_binaries(m, x, v, fraction: float, seed: int, rbar_pc: float = 1.0,
                 amin_au: float = 0.005, amax_au: float = 50.0):
    """Turn ``fraction`` of the particles into members of binaries.

    Consecutive particle pairs (2k, 2k+1) for k < fraction*n/2 share the centre
    of mass and velocity of particle 2k; masses are kept, the pair is placed on a
    Kepler orbit with log-uniform semi-major axis (AU, converted with the length
    scale ``rbar_pc``) and thermal eccentricity f(e) = 2e, at a random phase and
    orientation.  Returns new (m, x, v) arrays (the total mass is unchanged)."""
    rng = _rng(seed)
    m = m.copy(); x = x.copy(); v = v.copy()
    n = m.size
    nb = int(fraction * n) // 2
    if nb == 0:
        return m, x, v
    i1 = np.arange(nb) * 2
    i2 = i1 + 1
    mt = m[i1] + m[i2]
    # centre of mass of the pair = the old particle 2k's phase-space point
    xc = x[i1].copy(); vc = v[i1].copy()
    a = np.exp(rng.uniform(math.log(amin_au), math.log(amax_au), nb)) / (rbar_pc * AU_PER_PC)
    e = np.sqrt(rng.uniform(0.0, 1.0, nb))
    # eccentric anomaly from a uniform mean anomaly (Newton iterations)
    mean = rng.uniform(0.0, 2.0 * math.pi, nb)
    ea = mean.copy()
    for _ in range(30):
        ea = ea - (ea - e * np.sin(ea) - mean) / (1.0 - e * np.cos(ea))
    cosE, sinE = np.cos(ea), np.sin(ea)
    rrel = a * (1.0 - e * cosE)
    px = a * (cosE - e); py = a * np.sqrt(1.0 - e * e) * sinE
    nmot = np.sqrt(mt / a ** 3)
    vx = -a * nmot * sinE / (1.0 - e * cosE)
    vy = a * nmot * np.sqrt(1.0 - e * e) * cosE / (1.0 - e * cosE)
    # random orientation: orthonormal frame from two random unit vectors
    ez = _random_unit_vectors(rng, nb)
    t = _random_unit_vectors(rng, nb)
    ex = np.cross(ez, t)
    ex /= np.linalg.norm(ex, axis=1)[:, None]
    ey = np.cross(ez, ex)
    drel = px[:, None] * ex + py[:, None] * ey
    dvel = vx[:, None] * ex + vy[:, None] * ey
    f1 = (m[i2] / mt)[:, None]
    f2 = (m[i1] / mt)[:, None]
    x[i1] = xc + f1 * drel
    x[i2] = xc - f2 * drel
    v[i1] = vc + f1 * dvel
    v[i2] = vc - f2 * dvel
    del rrel
    return m, x, v


def rs_for_counts(x: np.ndarray, idx: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """Neighbour radius for each i in ``idx`` placed half-way between the
    counts[k]-th and (counts[k]+1)-th nearest particle (self included), so that
    the plain r < RS test selects exactly ``counts[k]`` particles."""
    from scipy.spatial import cKDTree
    tree = cKDTree(x)
    kmax = int(counts.max()) + 1
    out = np.empty(idx.size)
    step = max(1, 4_000_000 // max(kmax, 1))
    for s in range(0, idx.size, step):
        sl = slice(s, min(idx.size, s + step))
        d, _ = tree.query(x[idx[sl]], k=kmax)
        c = counts[sl]
        rows = np.arange(d.shape[0])
        out[sl] = 0.5 * (d[rows, c - 1] + d[rows, c])
    return out


def lognormal_counts(n: int, seed: int, median: float = 80.0, sigma: float = 0.5,
                     cap: int = 500, lo: int = 2) -> np.ndarray:
    rng = _rng(seed)
    c = np.exp(rng.normal(math.log(median), sigma, n))
    return np.clip(np.rint(c), lo, cap).astype(np.int64)


def precision_study(nj: int = 1 << 20, ni: int = 1 << 17, seed: int = 20261204):
    """The N = 1,048,576 precision-study system and its largest block-step i-set:
    King W0 = 6 with 50 % of the systems split into binaries, Kroupa masses, all nj
    as j-particles, ni i-particles drawn at random, RS giving log-normal neighbour
    counts (median 80, cap 500), dtr log-uniform in [2^-10, 2^-4].  Returns m, x, v,
    idx (sorted i indices), h2 = RS^2 and dtr per i."""
    mass = kroupa_masses(nj, seed)
    m, x, v = king(nj, seed + 1, w0=6.0, masses=mass)
    m, x, v = add_binaries(m, x, v, 0.5, seed + 2)
    rng = _rng(seed + 3)
    idx = np.sort(rng.choice(nj, size=ni, replace=False))
    counts = lognormal_counts(ni, seed + 4, median=80.0)
    rs = rs_for_counts(x, idx, counts)
    dtr = np.exp(rng.uniform(np.log(2.0 ** -10), np.log(2.0 ** -4), ni))
    return m, x, v, idx, rs * rs, dtr
