"""Pins the CPU restatement of the neighbour-radius rules and of the RS network
(oracle/rsnet.py): hand-computed rule cases, bfloat16 rounding cases, a
loop-by-loop forward pass and finite-difference gradients."""
import numpy as np
import pytest

import rsnet as R

GLOB = dict(rdens=np.zeros(3), rscale=1.0, rc2=0.04, rsfac=10.0, nnbopt=100, nnbmax=550,
            znbmin=5.5)


def one(rs, nnb, nnb0, x, v, dtr, **kw):
    g = dict(GLOB)
    g.update(kw)
    return float(R.regcor_rs(np.array([rs]), np.array([nnb]), np.array([nnb0]),
                             np.array([x], float), np.array([v], float), np.array([dtr]), **g)[0])


def test_initial_rs_and_step():
    rs, st = R.initial_rs(np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 4.0]]), 0.3, 0.125)
    assert np.allclose(rs, [0.3, 0.3 * np.sqrt(26.0)])
    assert np.allclose(st, [0.125, 0.125])            # capped at SMAX


def test_overflow_rescale_cases():
    # -NNB > NNBOPT: (NNBOPT/|NNB|)^0.333; else (NNBOPT/NNBMAX)^0.4; NB_FLAG=0: 0.9
    assert np.isclose(R.overflow_rs(1.0, -800, 100, 550), (100 / 800) ** 0.333)
    assert np.isclose(R.overflow_rs(2.0, -60, 100, 50), 2.0 * (100 / 50) ** 0.4)
    assert np.isclose(R.overflow_rs(2.0, -800, 100, 550, nb_flag=0), 1.8)


def test_rule_fixed_point_at_nnbopt():
    """NNB = NNB0 = NNBOPT inside the core: FAC = 1, volume ratio 1, RS unchanged."""
    assert one(0.05, 100, 100, [0.1, 0.0, 0.0], [0.0, 0.3, 0.0], 0.01) == 0.05


def test_rule_growth_hand_computed():
    """NNB = NNB0 = 50 in the core, large step (A5 = 0.25): FAC = 1.1, A4 = 1.1,
    A3 = 0.1/3, A1 = 1 + A3 - A3^2."""
    a3 = 0.1 / 3.0
    assert np.isclose(one(0.05, 50, 50, [0.1, 0.0, 0.0], [0.0, 0.3, 0.0], 0.1),
                      0.05 * (1.0 + a3 - a3 * a3))


def test_rule_shrink_with_inflow_hand_computed():
    """NNB = NNB0 = 400 > NBAVE = 250, moving inwards: A3 = min(370, 225), A4 - 1 =
    -0.4375 clamped to -0.25, A1 = 1 - 1/12 - 1/144, then x (1 - 150 / 1200)."""
    a3 = -0.25 / 3.0
    got = one(0.05, 400, 400, [0.1, 0.0, 0.0], [-0.3, 0.0, 0.0], 0.1)
    assert np.isclose(got, 0.05 * (1.0 + a3 - a3 * a3) * (1.0 - 150.0 / 1200.0))


def test_rule_small_step_limit():
    """The change is limited to RSFAC * DTR: NNB = 50 with DTR = 0.003, RSFAC = 10
    gives A4 - 1 = 0.03, A1 = 1.01 - 0.0001."""
    got = one(0.05, 50, 50, [0.1, 0.0, 0.0], [0.0, 0.3, 0.0], 0.003)
    assert np.isclose(got, 0.05 * (1.0 + 0.01 - 0.0001))


def test_rule_no_neighbours():
    """NNB = 0: outward motion shrinks to max(0.75 RS, 0.1 RSCALE), inward grows 1.1."""
    assert np.isclose(one(0.2, 0, 10, [2.0, 0.0, 0.0], [0.5, 0.0, 0.0], 0.1), 0.15)
    assert np.isclose(one(0.05, 0, 10, [2.0, 0.0, 0.0], [0.5, 0.0, 0.0], 0.1), 0.1)
    assert np.isclose(one(0.2, 0, 10, [2.0, 0.0, 0.0], [-0.5, 0.0, 0.0], 0.1), 0.22)


def test_rule_outer_membership_reduction():
    """Outside the half-mass radius the predicted membership is scaled by
    max(RSCALE / r, 0.1); with NNB = NNB0 = 100 at r = 2: A3 = 50, A4 = 0.5,
    radial factor (1 + RIDOT DTR / RI2) with RIDOT = 0, clamp to -0.25."""
    a3 = -0.25 / 3.0
    got = one(0.05, 100, 100, [2.0, 0.0, 0.0], [0.0, 0.3, 0.0], 0.1)
    assert np.isclose(got, 0.05 * (1.0 + a3 - a3 * a3))


def test_kth_radius_brute_force():
    rng = np.random.default_rng(3)
    xj, vj = rng.normal(size=(50, 3)), rng.normal(size=(50, 3))
    xi, vi = rng.normal(size=(2, 3)), rng.normal(size=(2, 3))
    dtr = np.array([0.0, 0.1])
    cand = [np.arange(50), np.arange(50)]
    got = R.kth_radius(xi, vi, dtr, xj, vj, cand, 7)
    for i in range(2):
        d = [min(np.linalg.norm(xj[j] - xi[i]), np.linalg.norm(xj[j] - xi[i] + dtr[i] * (vj[j] - vi[i])))
             for j in range(50)]
        assert np.isclose(got[i], sorted(d)[6])
        # exactly 7 within that radius (the k-th itself counts)
        assert sum(x <= got[i] for x in d) == 7


def test_bf16_rounding():
    f = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 3.0e-3],
                 np.float32)
    got = R.bf16(f)
    assert got[0] == 1.0
    assert got[1] == 1.0                      # tie -> even
    assert got[2] == 1.0 + 4 * 2 ** -8        # tie -> even (upwards)
    assert got[3] == 1.0 + 2 * 2 ** -8        # above the tie
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_forward_against_explicit_loops():
    p = R.init_params(11).astype(np.float64)
    p[R.OFF_B1:R.OFF_W2] = np.linspace(-0.2, 0.2, 64)
    p[R.OFF_B2:R.OFF_W3] = np.linspace(0.1, -0.1, 64)
    p[R.OFF_B3] = 0.3
    x = np.random.default_rng(2).normal(size=(3, 8)).astype(np.float32)
    y = R.forward(p, x, round_weights=False)
    for s in range(3):
        h1 = [max(0.0, p[R.OFF_B1 + k] + sum(p[R.OFF_W1 + 8 * k + f] * x[s, f] for f in range(8)))
              for k in range(64)]
        h2 = [max(0.0, p[R.OFF_B2 + o] + sum(p[R.OFF_W2 + 64 * o + k] * h1[k] for k in range(64)))
              for o in range(64)]
        ys = p[R.OFF_B3] + sum(p[R.OFF_W3 + o] * h2[o] for o in range(64))
        assert np.isclose(y[s], ys, rtol=1e-12, atol=1e-12)


def test_gradient_finite_differences():
    """Central differences of the FP64 loss (weights unrounded) along random
    parameter directions and along single coordinates of every block."""
    rng = np.random.default_rng(5)
    p = R.init_params(7).astype(np.float64)
    p[R.OFF_B1:R.OFF_W2] = rng.normal(size=64) * 0.1
    p[R.OFF_B2:R.OFF_W3] = rng.normal(size=64) * 0.1
    x = rng.normal(size=(40, 8))
    t = rng.normal(size=40)
    loss, g = R.loss_and_grad(p, x, t, round_weights=False)
    h = 1e-6
    for _ in range(6):
        d = rng.normal(size=R.NPARAM)
        lp, _ = R.loss_and_grad(p + h * d, x, t, round_weights=False)
        lm, _ = R.loss_and_grad(p - h * d, x, t, round_weights=False)
        assert np.isclose((lp - lm) / (2 * h), g @ d, rtol=1e-5, atol=1e-9)
    for k in (R.OFF_W1 + 3, R.OFF_B1 + 5, R.OFF_W2 + 700, R.OFF_B2 + 9, R.OFF_W3 + 2, R.OFF_B3):
        e = np.zeros(R.NPARAM)
        e[k] = 1.0
        lp, _ = R.loss_and_grad(p + h * e, x, t, round_weights=False)
        lm, _ = R.loss_and_grad(p - h * e, x, t, round_weights=False)
        assert np.isclose((lp - lm) / (2 * h), g[k], rtol=1e-5, atol=1e-9)
    assert np.isclose(loss, ((R.forward(p, x, round_weights=False) - t) ** 2).mean())


@pytest.mark.parametrize("n", [1, 63, 64, 65])
def test_gradient_is_a_mean(n):
    """Duplicating the batch leaves loss and gradient unchanged (1/n scaling)."""
    rng = np.random.default_rng(n)
    p = R.init_params(3)
    x = rng.normal(size=(n, 8))
    t = rng.normal(size=n)
    l1, g1 = R.loss_and_grad(p, x, t)
    l2, g2 = R.loss_and_grad(p, np.concatenate([x, x]), np.concatenate([t, t]))
    assert np.isclose(l1, l2) and np.allclose(g1, g2)


@pytest.mark.parametrize("nnbmax", [150, 550])
def test_network_aim_is_the_rules_fixed_point(nnbmax):
    """The RS network aims at NNBOPT aim^3 neighbours (nb6.rsdrive.net_aim): the count
    at which the rule leaves RS unchanged inside the core; fewer neighbours grow RS,
    more shrink it.  Short lists (NNBMAX = 150): 0.9 NBAVE = 90 < NNBOPT; long: NNBOPT."""
    from nb6.rsdrive import net_aim
    znbmin = max(0.01 * nnbmax, 1.0)
    k = int(round(100 * net_aim(100, nnbmax, znbmin) ** 3))
    assert k == (90 if nnbmax == 150 else 100)
    kw = dict(nnbmax=nnbmax, znbmin=znbmin)
    x, v = [0.1, 0.0, 0.0], [0.0, 0.3, 0.0]
    assert one(0.05, k, k, x, v, 0.01, **kw) == 0.05
    assert one(0.05, k - 20, k - 20, x, v, 0.01, **kw) > 0.05
    assert one(0.05, k + 20, k + 20, x, v, 0.01, **kw) < 0.05
