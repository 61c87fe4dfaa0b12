"""Device-resident j-particles with the GPU predictor (nb6_jp_*): the prediction
of xbpredall.F:55-59 on the GPU followed by gpunb_regf for blocks given by
address gives the same bits as predicting on the host, gpunb_send_ and
gpunb_regf_ with the predicted i-particles."""
import numpy as np
import pytest

import oracle as O
from helpers import rows_as_sets
from nb6 import ics, lib

pytestmark = pytest.mark.gpu


def host_predict(x0, v0, f, fdot, t0, time):
    s = (time - t0)[:, None]
    x = ((fdot * s + f) * s + v0) * s + x0
    v = (fdot * (1.5 * s) + f) * (2.0 * s) + v0
    return x, v


def state(n, seed):
    m, x0, v0 = ics.plummer(n, seed)
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, 3)) * 0.3
    fdot = rng.normal(size=(n, 3)) * 0.5
    t0 = rng.choice([0.0, 0.0625, 0.09375, 0.1171875], n)
    return m, x0, v0, f, fdot, t0


def test_device_prediction_equals_host_prediction():
    n, time = 8192, 0.125
    m, x0, v0, f, fdot, t0 = state(n, 41)
    addr = np.arange(1, n + 1, dtype=np.int32)
    blk = np.sort(np.random.default_rng(2).choice(n, 700, replace=False))
    xp, vp = host_predict(x0, v0, f, fdot, t0, time)
    rs = ics.rs_for_counts(xp, blk, np.full(blk.size, 40))
    h2, dtr = rs * rs, np.full(blk.size, 0.03)
    g = lib.Gpunb(n)
    jp = lib.JPred(n)
    try:
        g.send(m, xp, vp)
        host = g.regf(h2, dtr, xp[blk], vp[blk], 300, 250)
        # upload in two batches, in scrambled order, then predict on the device
        perm = np.random.default_rng(3).permutation(n)
        for part in np.array_split(perm, 2):
            jp.set(addr[part], x0[part], v0[part], f[part], fdot[part], m[part], t0[part])
        jp.predict_send(time, n)
        dev = jp.regf(blk + 1, h2, dtr, 300, 250)
        for a, b in zip(host[:3], dev[:3]):
            assert np.array_equal(a, b)
        assert rows_as_sets(host[3]) == rows_as_sets(dev[3])
        # correct a subset and predict to the next block time
        upd = np.random.default_rng(4).choice(n, 500, replace=False)
        x0[upd] += 1e-3
        f[upd] *= 0.5
        t0[upd] = time
        jp.set(addr[upd], x0[upd], v0[upd], f[upd], fdot[upd], m[upd], t0[upd])
        time2 = 0.140625
        xp2, vp2 = host_predict(x0, v0, f, fdot, t0, time2)
        g.send(m, xp2, vp2)
        host2 = g.regf(h2, dtr, xp2[blk], vp2[blk], 300, 250)
        jp.predict_send(time2, n)
        dev2 = jp.regf(blk + 1, h2, dtr, 300, 250)
        for a, b in zip(host2[:3], dev2[:3]):
            assert np.array_equal(a, b)
        assert rows_as_sets(host2[3]) == rows_as_sets(dev2[3])
    finally:
        jp.close()
        g.close()
    # and the lists are the FP64 oracle's at the predicted state
    ref = O.regf(m, xp2, vp2, h2, dtr, xp2[blk], vp2[blk], 300, 250)
    assert rows_as_sets(ref[3]) == rows_as_sets(dev2[3])


def test_bad_address_is_an_error():
    g = lib.Gpunb(64)
    jp = lib.JPred(64)
    L = lib.load()
    L.nb6_set_abort_on_error(0)
    try:
        with pytest.raises(RuntimeError):
            jp.set(np.array([0]), np.zeros((1, 3)), np.zeros((1, 3)), np.zeros((1, 3)),
                   np.zeros((1, 3)), np.ones(1), np.zeros(1))
        jp.predict_send(0.0, 10)
        with pytest.raises(RuntimeError):
            jp.regf(np.array([11]), np.ones(1), np.zeros(1), 10, 5)
    finally:
        L.nb6_set_abort_on_error(1)
        jp.close()
        g.close()
