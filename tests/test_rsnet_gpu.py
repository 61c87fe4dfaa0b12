"""Learned neighbour radius on the GPU: the network's forward pass and gradient
against the CPU oracle (oracle/rsnet.py, itself pinned by finite differences in
test_rsnet_cpu.py), training, the reference RS rule and the target radius
against the oracle, the fused prediction, and what the trained network gains
over the rule in a short run of regular block steps."""
import ctypes

import numpy as np
import pytest

import rsnet as R
from nb6 import ics, lib
from nb6.rsdrive import RSNet, Sim

pytestmark = pytest.mark.gpu


def dev(a):
    return lib.DeviceArray.from_host(np.ascontiguousarray(a))


@pytest.fixture
def net():
    n = RSNet(seed=7)
    yield n
    n.close()


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = R.init_params(seed)
    p[R.OFF_B1:R.OFF_W2] = rng.normal(size=64) * 0.2
    p[R.OFF_B2:R.OFF_W3] = rng.normal(size=64) * 0.2
    p[R.OFF_B3] = 0.1
    return p.astype(np.float32)


def test_forward_matches_oracle(net):
    p = random_params(3)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    mu = np.linspace(-1, 1, 8).astype(np.float32)
    sigma = np.linspace(0.5, 2.0, 8).astype(np.float32)
    net.set_norm(mu, sigma)
    feat = np.random.default_rng(4).normal(size=(1000, 8)).astype(np.float32) * 2
    fd, yd = dev(feat.T), lib.DeviceArray(1000 * 4)
    assert net.L.nb6_rsnet_forward_device(1000, fd.ptr, yd.ptr, None) == 0
    y = yd.to_host(np.float32, 1000)
    ref = R.forward(p, R.normalise(feat, mu, sigma))
    fd.free(); yd.free()
    err = np.abs(y - ref) / (np.abs(ref) + 1.0)
    print(f"forward max err {err.max():.2e}")
    assert err.max() < 1e-5


@pytest.mark.parametrize("n", [1, 100, 1000])
def test_gradient_matches_oracle(net, n):
    """GPU gradient (FP32 accumulation) against the FP64 oracle on the same
    bfloat16-rounded weights: relative error of the whole vector and of each
    parameter block."""
    p = random_params(5)
    net.set_params(p)
    mu = np.zeros(8, np.float32)
    sigma = np.ones(8, np.float32)
    net.set_norm(mu, sigma)
    rng = np.random.default_rng(n)
    feat = rng.normal(size=(n, 8)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    fd, td = dev(feat.T), dev(t)
    gd, ld = lib.DeviceArray(R.NPARAM * 4), lib.DeviceArray(4)
    assert net.L.nb6_rsnet_grad_device(n, fd.ptr, td.ptr, gd.ptr, ld.ptr, None) == 0
    g = gd.to_host(np.float32, R.NPARAM).astype(np.float64)
    loss = float(ld.to_host(np.float32, 1)[0])
    for a in (fd, td, gd, ld):
        a.free()
    rl, rg = R.loss_and_grad(p, feat, t)
    print(f"n={n}: loss {loss:.6f} vs {rl:.6f}; grad rel err {np.linalg.norm(g - rg) / np.linalg.norm(rg):.2e}")
    assert abs(loss - rl) <= 1e-5 * max(rl, 1e-3)
    assert np.linalg.norm(g - rg) <= 1e-5 * np.linalg.norm(rg)
    for a, b in ((R.OFF_W1, R.OFF_B1), (R.OFF_B1, R.OFF_W2), (R.OFF_W2, R.OFF_B2),
                 (R.OFF_B2, R.OFF_W3), (R.OFF_W3, R.OFF_B3), (R.OFF_B3, R.NPARAM)):
        assert np.linalg.norm(g[a:b] - rg[a:b]) <= 1e-4 * np.linalg.norm(rg[a:b]) + 1e-9


def test_training_fits_a_smooth_target(net):
    rng = np.random.default_rng(9)
    feat = rng.normal(size=(65536, 8)).astype(np.float32)
    t = (np.sin(feat[:, 0]) + 0.5 * feat[:, 1] * feat[:, 2] - 0.3 * feat[:, 7]).astype(np.float32)
    hist = net.train(feat, t, epochs=10, batch=2048, lr=3e-3)
    print("loss history", [round(h, 4) for h in hist])
    assert hist[-1] < 0.1 * np.var(t) and hist[-1] < hist[0]


def test_rule_kernel_matches_oracle():
    L = lib.load()
    rng = np.random.default_rng(11)
    n, nj, lmax = 96, 200, 8
    xj = rng.uniform(-3, 3, size=(nj, 3))
    vj = rng.normal(size=(nj, 3))
    sel = rng.choice(nj, n, replace=False)          # the block particles are j-set members
    x, v = xj[sel], vj[sel]
    nnb = rng.choice([0, 1, 2, 3, 20, 50, 100, 180, 400, 540], n).astype(np.int32)
    nnb0 = np.clip(nnb + rng.integers(-30, 30, n), 0, 550).astype(np.int32)
    dtr = np.exp(rng.uniform(np.log(1e-3), np.log(0.1), n))
    rs = rng.uniform(0.02, 0.5, n)
    lst = np.zeros((n, lmax), np.int32)
    nbl = {}
    for i in range(n):
        if nnb[i] <= 3:
            others = rng.choice(np.setdiff1d(np.arange(nj), [sel[i]]), nnb[i], replace=False)
            row = np.sort(np.r_[others, sel[i]])
            lst[i, 0] = row.size
            lst[i, 1:1 + row.size] = row
            nbl[i] = others
    glob = np.array([0.1, 0.0, 0.0, 1.0, 0.04, 10.0, 5.5])
    ref = R.regcor_rs(rs, nnb, nnb0, x, v, dtr, rdens=glob[:3], rscale=1.0, rc2=0.04, rsfac=10.0,
                      nnbopt=100, nnbmax=550, znbmin=5.5,
                      nb_positions=lambda i: (xj[nbl[i]], vj[nbl[i]]))
    bufs = [dev(np.concatenate([x.T, v.T])), dev(dtr), dev(nnb), dev(nnb0), dev(lst), dev(xj.T),
            dev(vj.T), dev(rs)]
    xv, dd, nd, n0, ld, xjd, vjd, rd = bufs
    assert L.nb6_rs_rule_device(n, n, xv.ptr, xv.ptr + 3 * n * 8, dd.ptr, nd.ptr, n0.ptr, ld.ptr, lmax,
                                xjd.ptr, vjd.ptr, nj, glob.ctypes.data, 100, 550, rd.ptr, None) == 0
    got = rd.to_host(np.float64, n)
    for b in bufs:
        b.free()
    assert np.allclose(got, ref, rtol=1e-13, atol=0)
    assert (got != rs).any()


def test_kth_radius_matches_oracle():
    """Target radius from a generous gpunb_regf list, against the oracle on the
    same candidates (FP64, same formula order: equal)."""
    m, x, v = ics.plummer(8192, 21)
    idx = np.arange(0, 8192, 7)
    rs = ics.rs_for_counts(x, idx, np.full(idx.size, 160))
    dtr = np.full(idx.size, 0.02)
    g = lib.Gpunb(8192)
    try:
        g.send(m, x, v)
        L = lib.load()
        ni, lmax, k = idx.size, 400, 101
        ib = np.concatenate([x[idx].T, v[idx].T, (rs * rs)[None], dtr[None]])
        di, do, dl, dk = dev(ib), lib.DeviceArray(7 * ni * 8), lib.DeviceArray(ni * lmax * 4), lib.DeviceArray(ni * 8)
        assert L.nb6_regf_device(ni, ni, di.ptr, lmax, lmax - 1, 0, do.ptr, dl.ptr, None) == 0
        assert L.nb6_rs_kth_device(ni, ni, di.ptr, dl.ptr, lmax, k, dk.ptr, None) == 0
        got = dk.to_host(np.float64, ni)
        lst = dl.to_host(np.int32, (ni, lmax))
        for a in (di, do, dl, dk):
            a.free()
    finally:
        g.close()
    ok = lst[:, 0] >= k
    assert ok.mean() > 0.9
    cand = [lst[i, 1:1 + lst[i, 0]] for i in np.nonzero(ok)[0]]
    ref = R.kth_radius(x[idx][ok], v[idx][ok], dtr[ok], x, v, cand, k)
    assert np.array_equal(got[ok], ref)
    assert (got[~ok] < 0).all()


def test_predict_fallback_and_h2(net):
    p = random_params(13)
    net.set_params(p)
    net.set_norm(np.zeros(8, np.float32), np.ones(8, np.float32))
    n = 300
    rng = np.random.default_rng(2)
    feat = rng.normal(size=(n, 8)).astype(np.float32)
    rs_prev = rng.uniform(0.05, 0.2, n)
    rs_rule = rng.uniform(0.05, 0.2, n)
    use = (rng.random(n) < 0.3).astype(np.int32)
    bufs = [dev(feat.T), dev(rs_prev), dev(rs_rule), dev(use), lib.DeviceArray(n * 8),
            dev(np.zeros((8, n)))]
    fd, pd, rd, ud, od, ibd = bufs
    assert net.L.nb6_rs_predict_device(n, fd.ptr, pd.ptr, rd.ptr, ud.ptr, od.ptr, ibd.ptr, n, 1,
                                       ctypes.c_double(0.25), None) == 0
    out = od.to_host(np.float64, n)
    ib = ibd.to_host(np.float64, (8, n))
    for b in bufs:
        b.free()
    y = np.clip(R.forward(p, feat), -4, 4)
    want = np.where(use == 1, rs_rule, rs_prev * np.exp(y))
    assert np.allclose(out, want, rtol=2e-5)
    assert np.array_equal(out[use == 1], rs_rule[use == 1])
    assert np.allclose(ib[6], out * out / 0.25, rtol=1e-15)


def test_features_closed_form():
    L = lib.load()
    n = 2
    x = np.array([[3.0, 0.0], [4.0, 0.0], [0.0, 1.0]])        # [3][n]
    acc = np.array([[0.0, 1e-3], [0.0, 0.0], [2.0, 0.0]])
    jrk = np.array([[10.0, 0.0], [0.0, 0.0], [0.0, 0.5]])
    mass = np.array([1e-3, 2e-6])
    rs = np.array([0.5, 0.1])
    dtr = np.array([2.0 ** -5, 2.0 ** -10])
    nnb = np.array([99, 0], np.int32)
    centre = np.array([0.0, 0.0, 1.0])
    bufs = [dev(x), dev(acc), dev(jrk), dev(mass), dev(rs), dev(dtr), dev(nnb), lib.DeviceArray(8 * n * 4)]
    assert L.nb6_rs_features_device(n, n, *[b.ptr for b in bufs[:7]], centre.ctypes.data,
                                    bufs[7].ptr, None) == 0
    f = bufs[7].to_host(np.float32, (8, n))
    for b in bufs:
        b.free()
    want0 = [np.log10(100 / (4 / 3 * np.pi * 0.125)), 2.0, np.log10(2.0), 1.0, np.log10(np.sqrt(26.0)),
             np.log10(0.5), -5 * np.log10(2), -3.0]
    assert np.allclose(f[:, 0], want0, rtol=1e-6)
    assert np.isclose(f[4, 1], np.log10(1e-6), rtol=1e-6)     # at the centre
    assert np.isclose(f[1, 1], 0.0)


@pytest.mark.parametrize("lmax", [300, 150])
def test_learned_radius_beats_the_rule_on_a_short_run(lmax):
    """King W0 = 7, N = 16384, NNBOPT = 50: pairs from regular block steps under
    the rule train the network; then, from the same start, the network (rule as
    fallback) keeps counts closer to NNBOPT than the rule does, and does not
    overflow more.  LMAX = 150 is a short list (NNBMAX = 100): both policies aim at
    0.9 NBAVE = 45 neighbours there."""
    n = 16384
    m, x, v = ics.king(n, 31, 7.0, masses=ics.kroupa_masses(n, 32))
    rec = []
    s = Sim(m, x, v, nnbopt=50, lmax=lmax)
    try:
        assert np.isclose(s.aim ** 3 * 50, 50 if lmax == 300 else 45)
        s.start()
        s.run(120, "rule", record=rec)
    finally:
        s.close()
    feat = np.concatenate([r[0] for r in rec])
    tgt = np.concatenate([r[1] for r in rec])
    assert feat.shape[0] > 20000
    net = RSNet(seed=3)
    try:
        hist = net.train(feat, tgt, epochs=8, batch=2048, lr=3e-3)
        print("train", len(tgt), "pairs, loss", [round(h, 4) for h in hist], "target var", float(tgt.var()))
        stats = {}
        for pol in ("rule", "net"):
            s = Sim(m, x, v, nnbopt=50, lmax=lmax)
            try:
                s.start(pol, net if pol == "net" else None)
                st = s.run(120, pol, net=net if pol == "net" else None)
                stats[pol] = dict(st)
            finally:
                s.close()
    finally:
        net.close()
    for pol, st in stats.items():
        print(pol, {k: st[k] for k in ("samples", "overflow_events", "miss30", "absdev", "fallback",
                                       "overflow_by_step")},
              "mean |ln(c/NNBOPT)|", st["absdev"] / st["samples"])
    r, q = stats["rule"], stats["net"]
    # on the short list ln(50/45) = 0.105 of each policy's mean is the shared aim itself
    # (measured: 0.118 against 0.164 there, 40 overflow re-runs of the rule against none)
    assert q["absdev"] / q["samples"] < (0.7 if lmax == 300 else 1.0) * r["absdev"] / r["samples"]
    assert q["miss30"] / q["samples"] < r["miss30"] / r["samples"]
    assert q["overflow_events"] <= r["overflow_events"]
