"""Checks of the FP64 oracle against what the mathematics fixes: closed forms,
finite differences, conservation laws, decompositions and invariances.  Each
check is chosen so that a plausible slip in the oracle (a dropped term, a wrong
sign or factor, a transposed operand, an off-by-one in the list) fails it."""
import numpy as np
import pytest

import oracle as O
from nb6 import ics

TINY_H2 = 1e-30      # only the self pair (r2 = 0) is a neighbour


def two_body(xj, vj, mj=0.3):
    m = np.array([0.7, mj])
    x = np.array([[0.0, 0.0, 0.0], xj], float)
    v = np.array([[0.1, -0.2, 0.05], vj], float)
    return m, x, v


def test_two_body_closed_form():
    m, x, v = two_body([0.3, -0.4, 1.2], [0.5, 0.1, -0.3])
    acc, jrk, pot, lst = O.regf(m, x, v, np.full(1, TINY_H2), np.zeros(1), x[:1], v[:1], 10, 5)
    d = x[1] - x[0]
    r = np.linalg.norm(d)
    assert np.allclose(acc[0], m[1] * d / r ** 3, rtol=1e-14)
    assert np.isclose(pot[0], m[1] / r, rtol=1e-14)
    assert lst[0, 0] == 1 and lst[0, 1] == 0          # only itself


def test_jerk_is_time_derivative_of_acc():
    """Central finite difference of acc along straight-line motion gives the jerk."""
    m, x, v = two_body([0.3, -0.4, 1.2], [0.5, 0.1, -0.3])
    h = 1e-5

    def acc_at(t):
        xt = x + t * v
        a, _, _, _ = O.regf(m, xt, v, np.full(1, TINY_H2), np.zeros(1), xt[:1], v[:1], 10, 5)
        return a[0]
    _, jrk, _, _ = O.regf(m, x, v, np.full(1, TINY_H2), np.zeros(1), x[:1], v[:1], 10, 5)
    fd = (acc_at(h) - acc_at(-h)) / (2 * h)
    assert np.allclose(jrk[0], fd, rtol=1e-8, atol=1e-12)


def test_ring_centre_cancels():
    n = 12
    ang = 2 * np.pi * np.arange(n) / n
    x = np.zeros((n + 1, 3))
    x[1:, 0], x[1:, 1] = 2.0 * np.cos(ang), 2.0 * np.sin(ang)
    m = np.full(n + 1, 0.1)
    v = np.zeros((n + 1, 3))
    acc, jrk, pot, _ = O.regf(m, x, v, np.full(1, TINY_H2), np.zeros(1), x[:1], v[:1], 20, 15)
    assert np.abs(acc[0]).max() < 1e-15 and np.abs(jrk[0]).max() < 1e-15
    assert np.isclose(pot[0], n * 0.1 / 2.0, rtol=1e-14)


def test_momentum_conservation_and_potential_energy():
    m, x, v = ics.plummer(512, 3, masses=ics.kroupa_masses(512, 4))
    n = m.size
    acc, jrk, pot, _ = O.regf(m, x, v, np.full(n, TINY_H2), np.zeros(n), x, v, 10, 5)
    scale = (m[:, None] * np.abs(acc)).sum()
    assert np.abs((m[:, None] * acc).sum(0)).max() < 1e-13 * scale
    assert np.abs((m[:, None] * jrk).sum(0)).max() < 1e-13 * (m[:, None] * np.abs(jrk)).sum()
    # the regular potential over all non-self pairs equals the all-pairs potential
    assert np.allclose(pot, O.pot(m, x), rtol=1e-13)


def test_neighbour_split_adds_up_to_the_direct_sum():
    """regular part (non-neighbours) + direct sum over the listed neighbours =
    full direct sum, for any RS."""
    m, x, v = ics.plummer(2048, 5)
    idx = np.arange(0, 2048, 37)
    rs = ics.rs_for_counts(x, idx, np.full(idx.size, 40))
    dtr = np.full(idx.size, 0.02)
    acc, jrk, pot, lst = O.regf(m, x, v, rs * rs, dtr, x[idx], v[idx], 200, 150)
    facc, fjrk, fpot, _ = O.regf(m, x, v, np.full(idx.size, TINY_H2), dtr, x[idx], v[idx], 200, 150)
    for k, i in enumerate(idx):
        nb = [j for j in lst[k, 1:1 + lst[k, 0]] if j != i]
        sub = np.r_[i, nb]
        a1, j1, p1, _ = O.regf(m[sub], x[sub], v[sub], np.full(1, TINY_H2), dtr[k:k + 1],
                               x[i:i + 1], v[i:i + 1], 200, 150)
        assert np.allclose(acc[k] + a1[0], facc[k], rtol=1e-12, atol=1e-14)
        assert np.allclose(jrk[k] + j1[0], fjrk[k], rtol=1e-11, atol=1e-13)
        assert np.isclose(pot[k] + p1[0], fpot[k], rtol=1e-13)


def test_neighbour_lists_ascending_and_sized():
    m, x, v = ics.plummer(4096, 6)
    idx = np.arange(0, 4096, 9)
    rs = ics.rs_for_counts(x, idx, np.full(idx.size, 60))
    _, _, _, lst = O.regf(m, x, v, rs * rs, np.zeros(idx.size), x[idx], v[idx], 200, 150)
    for k, i in enumerate(idx):
        row = lst[k, 1:1 + lst[k, 0]]
        assert lst[k, 0] == 60                   # dtr = 0: exactly the 60 inside RS
        assert (np.diff(row) > 0).all() and i in row


def test_velocity_criterion():
    """A body outside RS that will be inside after dtr is a neighbour; the same
    body receding is not (gpunb.velocity.cu:178-188)."""
    m = np.array([1e-3, 1e-3, 1e-3])
    x = np.array([[0, 0, 0], [1.0, 0, 0], [-1.0, 0, 0]], float)
    v = np.array([[0, 0, 0], [-10.0, 0, 0], [-10.0, 0, 0]], float)
    _, _, _, lst = O.regf(m, x, v, np.full(1, 0.5 ** 2), np.full(1, 0.1), x[:1], v[:1], 10, 5)
    assert sorted(lst[0, 1:1 + lst[0, 0]]) == [0, 1]


def test_mass_criterion_is_rescaled_radius():
    """With equal masses m0, m_flag=1 and h2 = H equals m_flag=0 and h2 = m0 H."""
    m, x, v = ics.plummer(2048, 7)
    idx = np.arange(0, 2048, 17)
    H = np.full(idx.size, 0.01)
    dtr = np.full(idx.size, 0.01)
    a1, j1, p1, l1 = O.regf(m, x, v, H / m[0], dtr, x[idx], v[idx], 300, 250, 1)
    a0, j0, p0, l0 = O.regf(m, x, v, H, dtr, x[idx], v[idx], 300, 250, 0)
    assert np.array_equal(l0[:, 0], l1[:, 0])
    for k in range(idx.size):
        assert np.array_equal(l0[k, 1:1 + l0[k, 0]], l1[k, 1:1 + l1[k, 0]])
    assert np.array_equal(a0, a1) and np.array_equal(p0, p1)


def test_j_permutation_invariance():
    m, x, v = ics.plummer(1024, 8)
    idx = np.arange(0, 1024, 11)
    rs = ics.rs_for_counts(x, idx, np.full(idx.size, 30))
    dtr = np.full(idx.size, 0.01)
    perm = np.random.default_rng(9).permutation(1024)
    inv = np.argsort(perm)
    a, j, p, l = O.regf(m, x, v, rs * rs, dtr, x[idx], v[idx], 100, 80)
    ap, jp, pp, lp = O.regf(m[perm], x[perm], v[perm], rs * rs, dtr, x[idx], v[idx], 100, 80)
    assert np.allclose(a, ap, rtol=1e-12) and np.allclose(p, pp, rtol=1e-13)
    for k in range(idx.size):
        assert sorted(perm[lp[k, 1:1 + lp[k, 0]]]) == list(l[k, 1:1 + l[k, 0]])
    del inv


def test_overflow_returns_minus_exact_count():
    m, x, v = ics.plummer(2048, 10)
    idx = np.arange(0, 2048, 64)
    counts = np.where(np.arange(idx.size) % 2 == 0, 30, 90)
    rs = ics.rs_for_counts(x, idx, counts)
    _, _, _, lst = O.regf(m, x, v, rs * rs, np.zeros(idx.size), x[idx], v[idx], 100, 64)
    assert np.array_equal(lst[:, 0], np.where(counts > 64, -counts, counts))


def test_firr_closed_form_two_body_with_prediction():
    """Irregular force of a pair whose members move on known third-order
    trajectories from different times t0."""
    x = np.array([[0.0, 0.0, 0.0], [0.2, 0.1, -0.1]])
    v = np.array([[0.1, 0.0, 0.0], [-0.1, 0.2, 0.0]])
    a2 = np.array([[0.01, 0.0, 0.02], [0.0, -0.03, 0.0]])
    j6 = np.array([[0.0, 0.001, 0.0], [0.002, 0.0, 0.0]])
    m = np.array([0.4, 0.6])
    t0 = np.array([0.25, 0.5])
    ti = 0.75
    lists = np.zeros((2, 4), np.int32)
    lists[0, :2] = [1, 2]
    lists[1, :2] = [1, 1]
    acc, jrk, nn = O.firr(ti, np.array([1, 2]), x, v, a2, j6, m, t0, lists, 4)

    def state(k):
        s = ti - t0[k]
        xp = x[k] + v[k] * s + a2[k] * s ** 2 + j6[k] * s ** 3
        vp = v[k] + 2 * a2[k] * s + 3 * j6[k] * s ** 2
        return xp, vp
    for k, o in ((0, 1), (1, 0)):
        xi, vi = state(k)
        xj, vj = state(o)
        d, w = xj - xi, vj - vi
        r = np.linalg.norm(d)
        assert np.allclose(acc[k], m[o] * d / r ** 3, rtol=1e-13)
        assert np.allclose(jrk[k], m[o] * (w / r ** 3 - 3 * d.dot(w) * d / r ** 5), rtol=1e-12)
        assert nn[k] == o + 1


def test_firr_plus_regf_is_the_full_force():
    """At ti = t0 (no prediction) the neighbour force over the regf lists plus the
    regular force is the full direct force."""
    m, x, v = ics.plummer(1024, 12)
    n = m.size
    rs = ics.rs_for_counts(x, np.arange(n), np.full(n, 40))
    dtr = np.zeros(n)
    acc, jrk, pot, lst = O.regf(m, x, v, rs * rs, dtr, x, v, 100, 80)
    facc, fjrk, _, _ = O.regf(m, x, v, np.full(n, TINY_H2), dtr, x, v, 100, 80)
    lists = np.zeros((n, 100), np.int32)
    for i in range(n):
        row = [j + 1 for j in lst[i, 1:1 + lst[i, 0]] if j != i]
        lists[i, 0] = len(row)
        lists[i, 1:1 + len(row)] = row
    z = np.zeros((n, 3))
    iacc, ijrk, nn = O.firr(0.5, np.arange(1, n + 1), x, v, z, z, m, np.full(n, 0.5), lists, 100)
    assert np.allclose(acc + iacc, facc, rtol=1e-11, atol=1e-13)
    assert np.allclose(jrk + ijrk, fjrk, rtol=1e-10, atol=1e-12)
    # nearest neighbour = argmin distance over the list
    for i in range(0, n, 7):
        row = lists[i, 1:1 + lists[i, 0]] - 1
        d2 = ((x[row] - x[i]) ** 2).sum(1)
        assert nn[i] == row[np.argmin(d2)] + 1


@pytest.mark.parametrize("seed", [1, 2])
def test_plummer_potential_profile(seed):
    """At large N the all-pairs potential follows phi(r) = M / sqrt(r^2 + a^2)
    with a = 3 pi / 16 in N-body units (statistically)."""
    m, x, v = ics.plummer(4096, seed)
    p = O.pot(m, x)
    a = 3 * np.pi / 16
    r = np.linalg.norm(x, axis=1)
    pred = 1.0 / np.sqrt(r * r + a * a)
    inner = r < 2.0
    assert np.median(np.abs(p[inner] / pred[inner] - 1)) < 0.03
