"""GPU parity of irr_simd_firr_vec_ (and the gpuirr_* aliases) against the FP64
oracle."""
import numpy as np
import pytest

import oracle as O
from helpers import c1_inputs
from nb6 import lib

pytestmark = pytest.mark.gpu

# The irregular force keeps the reference's record format (irr.avx.cpp:57-233:
# positions FP64, velocity, a/2 and adot/6 FP32) and FP32 pair arithmetic, so its
# jerk carries the FP32 rounding of the velocities (up to about 6e-6 measured where
# the neighbour terms cancel); acc stays within 2e-6.
REL_TOL = 2.0e-6
JRK_TOL = 1.0e-5


def irr_state(n=1024, seed=5):
    """Particles with predictor records, per-particle times on a block grid,
    neighbour lists from the FP64 regular-force oracle (self removed, 1-based)."""
    m, x, v, h2, dtr = c1_inputs(n=n)
    rng = np.random.Generator(np.random.PCG64(seed))
    a2 = 0.5 * rng.normal(0, 0.3, (n, 3))
    j6 = rng.normal(0, 0.2, (n, 3)) / 6.0
    t0 = np.floor(rng.uniform(0, 64, n)) / 1024.0
    _, _, _, nb = O.regf(m, x, v, h2, dtr, x, v, 400, 350)
    lists = np.zeros((n, 400), np.int32)
    for i in range(n):
        c = nb[i, 0]
        row = [j + 1 for j in nb[i, 1:1 + c] if j != i]
        lists[i, 0] = len(row)
        lists[i, 1:1 + len(row)] = row
    return m, x, v, a2, j6, t0, lists


@pytest.mark.parametrize("alias", [False, True])
def test_firr_matches_fp64(alias):
    m, x, v, a2, j6, t0, lists = irr_state()
    n = m.size
    ti = 64.0 / 1024.0
    irr = lib.IrrSimd(n, 400, alias=alias)
    for i in range(n):
        irr.set_jp(i + 1, x[i], v[i], a2[i], j6[i], m[i], t0[i])
        irr.set_list(i + 1, lists[i])
    addr = np.arange(1, n + 1, dtype=np.int32)[::3]
    acc, jrk, nn = irr.firr_vec(ti, addr)
    racc, rjrk, rnn = O.firr(ti, addr, x, v, a2, j6, m, t0, lists, 400)
    irr.profile()
    irr.close()
    ea = np.linalg.norm(acc - racc, axis=1) / np.linalg.norm(racc, axis=1)
    ej = np.linalg.norm(jrk - rjrk, axis=1) / np.linalg.norm(rjrk, axis=1)
    print(f"irr: acc max {ea.max():.3e} jrk max {ej.max():.3e} nnbid agree "
          f"{(nn == rnn).mean():.4f}", flush=True)
    assert ea.max() <= REL_TOL and ej.max() <= JRK_TOL
    assert (nn == rnn).all()


def test_firr_update_one_particle_and_list():
    """set_jp / set_list after a firr call take effect at the next call; a
    repeated set_list of one particle keeps the last list."""
    m, x, v, a2, j6, t0, lists = irr_state(n=512, seed=9)
    n = m.size
    irr = lib.IrrSimd(n, 400)
    for i in range(n):
        irr.set_jp(i + 1, x[i], v[i], a2[i], j6[i], m[i], t0[i])
        irr.set_list(i + 1, lists[i])
    ti = 0.0625
    addr = np.array([1, 2, 3, 100], np.int32)
    irr.firr_vec(ti, addr)
    x2 = x.copy(); x2[5] += 1e-3
    irr.set_jp(6, x2[5], v[5], a2[5], j6[5], m[5], t0[5])
    lists2 = lists.copy()
    short = lists[2, 1:1 + lists[2, 0]][: 3]
    irr.set_list(3, np.r_[7, lists[2, 1:8]].astype(np.int32))
    lists2[2, 0] = len(short); lists2[2, 1:1 + len(short)] = short
    irr.set_list(3, np.r_[len(short), short].astype(np.int32))
    acc, jrk, nn = irr.firr_vec(ti, addr)
    irr.close()
    racc, rjrk, rnn = O.firr(ti, addr, x2, v, a2, j6, m, t0, lists2, 400)
    ea = np.linalg.norm(acc - racc, axis=1) / np.linalg.norm(racc, axis=1)
    assert ea.max() <= REL_TOL
    assert (nn == rnn).all()


def test_firr_golden_inputs():
    """Reference AVX irregular force on its golden inputs: our error against the
    FP64 oracle is within REL_TOL and nearest neighbours agree with FP64."""
    from pathlib import Path
    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "irr_plummer1k.npz"))
    n = g["m"].size
    irr = lib.IrrSimd(n, 400)
    for i in range(n):
        irr.set_jp(i + 1, g["x"][i], g["v"][i], g["a2"][i], g["j6"][i], g["m"][i], g["t0"][i])
        irr.set_list(i + 1, g["lists"][i])
    acc, jrk, nn = irr.firr_vec(float(g["ti"]), g["addr"])
    irr.close()
    racc, rjrk, rnn = O.firr(float(g["ti"]), g["addr"], g["x"], g["v"], g["a2"], g["j6"], g["m"],
                             g["t0"], g["lists"], 400)
    ea = np.linalg.norm(acc - racc, axis=1) / np.linalg.norm(racc, axis=1)
    e_avx = np.linalg.norm(g["acc"] - racc, axis=1) / np.linalg.norm(racc, axis=1)
    assert ea.max() <= REL_TOL and ea.max() <= e_avx.max()
    assert (nn == rnn).all()


def test_firr_fed_from_device_equals_host_path():
    """Predictor records from the nb6_jp_* store and lists straight from the device
    gpunb_regf rows give the same bits as irr_simd_set_jp_ / irr_simd_set_list_
    with the host lists (util_gpu.F:101-110: +IFIRST, self dropped) and
    irr_simd_firr_vec_."""
    m, x, v, a2, j6, t0, _ = irr_state(n=2048, seed=8)
    n = m.size
    ti = 64.0 / 1024.0
    rng = np.random.default_rng(9)
    blk = np.sort(rng.choice(n, 600, replace=False))
    from nb6 import ics
    rs = ics.rs_for_counts(x, blk, np.full(blk.size, 30))
    h2, dtr = rs * rs, np.full(blk.size, 0.02)
    addr = np.arange(1, n + 1, dtype=np.int32)
    L = lib.load()
    g = lib.Gpunb(n)
    jp = lib.JPred(n)
    irr = lib.IrrSimd(n, 400)
    try:
        jp.set(addr, x, v, a2, j6, m, t0)
        jp.predict_send(ti, n)
        acc, jrk, pot, lst = jp.regf(blk + 1, h2, dtr, 400, 350)
        # host path
        for i in range(n):
            irr.set_jp(i + 1, x[i], v[i], a2[i], j6[i], m[i], t0[i])
        for k, i in enumerate(blk):
            c = lst[k, 0]
            row = [j + 1 for j in lst[k, 1:1 + c] if j != i]
            irr.set_list(i + 1, np.r_[len(row), row])
        ha, hj, hn = irr.firr_vec(ti, blk + 1)
        # device path: overwrite the records and lists from the device, same call
        L.nb6_irr_sync_jp_(lib._ri(n))
        lib.check_error()
        ib = np.ascontiguousarray(np.concatenate([np.zeros((6, blk.size)), h2[None], dtr[None]]))
        da = lib.DeviceArray.from_host((blk + 1).astype(np.int32))
        dl = lib.DeviceArray.from_host(np.ascontiguousarray(lst))
        do = lib.DeviceArray(blk.size * 8 * 8)
        assert L.nb6_irr_set_lists_device(blk.size, da.ptr, dl.ptr, 400, 1, None) == 0
        assert L.nb6_irr_firr_device(ti, blk.size, da.ptr, do.ptr, None) == 0
        out = do.to_host(np.float64, (blk.size, 8))
        for d in (da, dl, do):
            d.free()
        del ib
    finally:
        irr.close()
        jp.close()
        g.close()
    assert np.array_equal(out[:, 0:3], ha) and np.array_equal(out[:, 3:6], hj)
    assert np.array_equal(out[:, 6].astype(int), hn)
    # the device-fed result against the FP64 oracle on the same inputs: the device
    # regf rows of the block (+IFIRST, self dropped), the predictor records
    lists = np.zeros((n, 400), np.int32)
    for k, i in enumerate(blk):
        row = [j + 1 for j in lst[k, 1:1 + lst[k, 0]] if j != i]
        lists[i, 0] = len(row)
        lists[i, 1:1 + len(row)] = row
    racc, rjrk, rnn = O.firr(ti, blk + 1, x, v, a2, j6, m, t0, lists, 400)
    ea = np.linalg.norm(out[:, 0:3] - racc, axis=1) / np.linalg.norm(racc, axis=1)
    ej = np.linalg.norm(out[:, 3:6] - rjrk, axis=1) / np.linalg.norm(rjrk, axis=1)
    print(f"device-fed irr vs FP64: acc {ea.max():.2e} jrk {ej.max():.2e}", flush=True)
    assert ea.max() <= REL_TOL and ej.max() <= JRK_TOL
    assert np.array_equal(out[:, 6].astype(int), rnn)


@pytest.mark.parametrize("lpi", [64, 16])
def test_firr_predict_modes_give_the_same_bits(lpi):
    """Records predicted per gathered neighbour (mode 0) and all records predicted once
    per call (mode 2) give identical results, and both match FP64."""
    m, x, v, a2, j6, t0, lists = irr_state(n=2048, seed=12)
    n = m.size
    ti = 64.0 / 1024.0
    L = lib.load()
    irr = lib.IrrSimd(n, 400)
    outs = []
    try:
        for i in range(n):
            irr.set_jp(i + 1, x[i], v[i], a2[i], j6[i], m[i], t0[i])
            irr.set_list(i + 1, lists[i, :1 + lists[i, 0]])
        addr = np.arange(1, n + 1, 3, dtype=np.int32)
        assert L.nb6_set_irr_config(lpi, 1) == 0
        for mode in (0, 2, 1):
            assert L.nb6_set_irr_predict(mode) == 0
            outs.append(irr.firr_vec(ti, addr))
    finally:
        L.nb6_set_irr_predict(1)
        L.nb6_set_irr_config(64, 1)
        irr.close()
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)
    racc, rjrk, rnn = O.firr(ti, addr, x, v, a2, j6, m, t0, lists, 400)
    err = np.linalg.norm(outs[0][0] - racc, axis=1) / np.linalg.norm(racc, axis=1)
    assert err.max() <= REL_TOL and np.array_equal(outs[0][2], rnn)
