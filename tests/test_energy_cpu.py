"""Pins the FP64 restatement of the energy path (oracle/energy.py: energy.F's GPU
branch with phicor.f) with direct sums over the physical bodies: without pairs POT
is the pair sum; with every c.m. listed and no separation limit, PHICOR turns each
c.m. interaction into the two component interactions, so PHIDBL and POT equal direct
sums over singles and components (the internal pair energy stays in EBIN)."""
import numpy as np

import energy as E
from helpers import energy_case


def _direct(x, m, i, js):
    d = x[js] - x[i]
    return float((m[js] / np.sqrt((d * d).sum(1))).sum())


def test_no_pairs_is_the_pair_sum_and_plummer_energy():
    x, v, m, lst, n, P, ht = energy_case(seed=3, nsing=400, npairs=0)
    zkin, pot, ebin, phi = E.energy(x, v, m, lst, n, 0, thr=1.0)
    w = sum(m[i] * m[j] / np.linalg.norm(x[i] - x[j]) for i in range(n) for j in range(i + 1, n))
    assert np.isclose(pot, w, rtol=1e-12)
    assert np.isclose(zkin, 0.5 * (m * (v * v).sum(1)).sum(), rtol=1e-14)
    assert ebin == 0.0
    assert abs((zkin - pot) + 0.25) < 0.01          # Plummer scaled to E = -1/4 (sampling noise)


def test_phicor_recovers_component_interactions():
    x, v, m, lst, n, P, ht = energy_case(seed=5, nsing=150, npairs=12, full=True)
    zkin, pot, ebin, phi = E.energy(x, v, m, lst, n, P, thr=1e30, ht=ht)
    phys = np.arange(n)                               # components and singles, 0-based
    for i in range(2 * P, n):                          # singles: all other physical bodies
        want = _direct(x, m, i, phys[phys != i])
        assert np.isclose(phi[i], want, rtol=1e-12), i
    for p in range(P):                                 # components: all but the partner
        for c, partner in ((2 * p, 2 * p + 1), (2 * p + 1, 2 * p)):
            js = phys[(phys != c) & (phys != partner)]
            assert np.isclose(phi[c], _direct(x, m, c, js), rtol=1e-12), c
    w = 0.0
    for i in range(n):
        for j in range(i + 1, n):
            if i < 2 * P and j == i + 1 and i % 2 == 0:
                continue                               # internal pair energy -> EBIN
            w += m[i] * m[j] / np.linalg.norm(x[i] - x[j])
    assert np.isclose(pot, w, rtol=1e-12)
    want_ebin = sum(m[2 * p] * m[2 * p + 1] * ht[p] / m[n + p] for p in range(P))
    assert np.isclose(ebin, want_ebin, rtol=1e-14)


def test_phicor_separation_limit():
    """CMSEP2 RMIN2 = 0: singles keep the c.m. point-mass potential; a pair's components
    still exchange each perturber's c.m. distance for their own (phicor.f:44-50)."""
    x, v, m, lst, n, P, ht = energy_case(seed=7, nsing=150, npairs=10, full=True)
    phi0 = E.gpupot_fp64(x, m, 2 * P + 1, n + P)
    zkin, pot, ebin, phi = E.energy(x, v, m, lst, n, P, thr=0.0, ht=ht)
    assert np.allclose(phi[2 * P:n], phi0[:n - 2 * P], rtol=1e-15, atol=0)
    for p in range(P):
        I = n + p                                      # 0-based c.m.
        js = lst[2 * p, 1:1 + lst[2 * p, 0]] - 1
        for c in (2 * p, 2 * p + 1):
            want = phi0[I - 2 * P] - _direct(x, m, I, js) + _direct(x, m, c, js)
            assert np.isclose(phi[c], want, rtol=1e-12)


def test_ghost_pairs_are_skipped_in_ebin():
    x, v, m, lst, n, P, ht = energy_case(seed=9, nsing=100, npairs=4)
    m = m.copy()
    m[n + 1] = 0.0
    zkin, pot, ebin, phi = E.energy(x, v, m, lst, n, P, thr=1e-4, ht=ht)
    want = sum(m[2 * p] * m[2 * p + 1] * ht[p] / m[n + p] for p in range(P) if m[n + p] > 0)
    assert np.isclose(ebin, want, rtol=1e-14)
