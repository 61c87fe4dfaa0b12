"""GPU regular corrector (nb6_regcor_device) against the oracle's restatement of
regcor_gpu.F + fpcorr.F: bit-identical state, lists and counters on blocks that
exercise every branch (losses, gains, SMIN retention, empty old and new lists,
NNB <= 3, c.m. bodies, KZ(38) = 0..3, NB_FLAG, M_FLAG, TIDAL(4), CHECKL's
high-velocity intruders with KZ(37) = 1 for single and c.m. bodies, KS components in
FPCORR), fed by CPU-built
rows or straight from nb6_regf_device, and on a sample of a 32768-particle block of a
262,144-body system."""
import numpy as np
import pytest

import regcor as R
from helpers import copy_state, regcor_case
from nb6 import ics, lib

pytestmark = pytest.mark.gpu

FIELDS = lib.REG3 + lib.REG1


def _gpu(S, P, addr, d_force, force_stride, d_rows, ni):
    rc = lib.Regcor(S)
    d_addr = lib.DeviceArray.from_host(np.asarray(addr, np.int32))
    d_cnt = lib.DeviceArray.from_host(np.zeros(lib.NCOUNT, np.int32))
    try:
        rc.correct(P, ni, d_addr, d_force, force_stride, d_rows, d_cnt)
        assert lib.load().nb6_sync() == 0
        out = rc.download()
        cnt = d_cnt.to_host(np.int32, (lib.NCOUNT,))
    finally:
        rc.free()
        d_addr.free()
        d_cnt.free()
    return out, cnt


def _counters(C):
    return [C["nbvoid"], C["nbflux"], C["nbsmin"], C["nbcorr"], C["unsupported"], C["nbfast"],
            C["ncmder"]]


def _same(got, want, label, idx=None):
    for k in FIELDS:
        a, b = got[k], want[k]
        if idx is not None:
            a, b = a[..., idx], b[..., idx]
        bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        assert not bad.any(), f"{label}: {k} differs at {np.argwhere(bad)[:5].tolist()} " \
                              f"({bad.sum()} values)"
    la, lb = got["list"], want["list"]
    if idx is not None:
        la, lb = la[idx], lb[idx]
    assert np.array_equal(la, lb), f"{label}: LIST differs"


CASES = [
    dict(seed=7, kz38=1),
    dict(seed=8, kz38=2),
    dict(seed=9, kz38=3),
    dict(seed=10, kz38=0, nb_flag=0, tidal4=0.37),
    dict(seed=11, kz38=1, m_flag=1),
    dict(seed=12, kz38=1, ncm=60, ni=400),
    dict(seed=13, kz38=2, frac_void=0.3, frac_empty_old=0.3, frac_few=0.2),
    dict(seed=14, kz38=1, ncm=50, ifirst=101, ni=350),
    dict(seed=15, kz38=1, kz37=1, nfast=40),
    dict(seed=16, kz38=1, kz37=1, nfast=40, ncm=60, ni=400, ks_old=0.5, ifirst=41),
    dict(seed=18, kz38=2, kz37=1, nfast=120, ncm=80, ni=500, ifirst=61, ks_old=0.3,
         frac_small_step=0.4),
]


@pytest.mark.parametrize("pack", [1, 2], ids=["packed", "direct"])
@pytest.mark.parametrize("kw", CASES, ids=lambda kw: "-".join(f"{k}{v}" for k, v in kw.items()))
def test_regcor_matches_oracle_bitwise(kw, pack):
    """Both ways of reading the neighbours' state (packed records, state arrays)."""
    S, P, addr, force, rows = regcor_case(**kw)
    print(f"regcor case {kw} ni={len(addr)} pack={pack}", flush=True)
    So = copy_state(S)
    rows_o = rows.copy()
    C = R.regcor_block(So, P, addr, force, rows_o)
    assert C["unsupported"] == 0
    if kw.get("kz37"):
        assert C["nbfast"] > 0
    if kw.get("ks_old"):
        assert C["ncmder"] > 0
    d_force = lib.DeviceArray.from_host(np.ascontiguousarray(force))
    d_rows = lib.DeviceArray.from_host(rows.astype(np.int32))
    lib.load().nb6_set_regcor_pack(pack)
    try:
        got, cnt = _gpu(S, P, addr, d_force, force.shape[1], d_rows, len(addr))
        rows_g = d_rows.to_host(np.int32, rows.shape)
    finally:
        lib.load().nb6_set_regcor_pack(0)
        d_force.free()
        d_rows.free()
    _same(got, So, str(kw))
    assert np.array_equal(rows_g, rows_o), "converted rows differ"
    assert cnt.tolist() == _counters(C)


def test_regcor_fed_by_device_regf():
    """regf output (force, rows) goes from nb6_regf_device to nb6_regcor_device without
    leaving the GPU; the oracle gets copies of the same force and rows."""
    S, P, addr, _, _ = regcor_case(seed=17, n=2048, ni=700, lmax=200, nnbopt=30)
    n, ni, lmax = len(S["body"]), len(addr), S["list"].shape[1]
    ia = addr - 1
    g = lib.Gpunb(n)
    try:
        g.send(S["body"], S["x"].T.copy(), S["xdot"].T.copy())
        dtr = P["time"] - S["t0r"][ia]
        ib = np.concatenate([S["x"][:, ia], S["xdot"][:, ia], (S["rs"][ia] ** 2)[None],
                             dtr[None]]).astype(np.float64)
        di = lib.DeviceArray.from_host(ib)
        do = lib.DeviceArray(7 * ni * 8)
        dl = lib.DeviceArray(ni * lmax * 4)
        rc = lib.load().nb6_regf_device(ni, ni, di.ptr, lmax, P["nnbmax"], 0, do.ptr, dl.ptr, None)
        assert rc == 0
        force = do.to_host(np.float64, (7, ni))
        rows = dl.to_host(np.int32, (ni, lmax))
        assert rows[:, 0].min() >= 0
        So = copy_state(S)
        rows_o = rows.copy()
        C = R.regcor_block(So, P, addr, force, rows_o)
        got, cnt = _gpu(S, P, addr, do, ni, dl, ni)
        rows_g = dl.to_host(np.int32, (ni, lmax))
        for d in (di, do, dl):
            d.free()
    finally:
        g.close()
    assert C["nbflux"] > 0 and C["nbsmin"] > 0
    _same(got, So, "regf-fed")
    assert np.array_equal(rows_g, rows_o)
    assert cnt.tolist() == _counters(C)


def test_regcor_large_block_sample():
    """N = 262,144 King W0 = 6 with binaries, a block of 32,768: regf (device) builds
    the new rows and the force, a second regf at radii one step back builds the old
    lists; 256 sampled particles are checked against the oracle bit for bit, and the
    counters of the whole block against a host recount of list changes."""
    n, ni, lmax, nnbmax = 1 << 18, 1 << 15, 600, 550
    m, x, v = ics.king(n, 20261204, 6.0)
    m, x, v = ics.add_binaries(m, x, v, 0.5, 20261204)
    rng = np.random.default_rng(5)
    addr0 = np.sort(rng.choice(n, ni, replace=False))
    counts = ics.lognormal_counts(ni, 3, median=60, cap=300)
    rs = np.full(n, 0.05)
    rs[addr0] = ics.rs_for_counts(x, addr0, counts)
    # regular steps: powers of two that keep |v| dtr below 0.2 RS (binaries are fast)
    vabs = np.sqrt((v * v).sum(1)) + 1e-30
    dtr_all = 2.0 ** np.clip(np.floor(np.log2(0.2 * rs / vabs)), -14, -3)
    time = 1.0
    g = lib.Gpunb(n)
    try:
        g.send(m, x, v)

        def regf(rad):
            ib = np.concatenate([x[addr0].T, v[addr0].T, (rad ** 2)[None],
                                 dtr_all[addr0][None]]).astype(np.float64)
            di = lib.DeviceArray.from_host(ib)
            do = lib.DeviceArray(7 * ni * 8)
            dl = lib.DeviceArray(ni * lmax * 4)
            assert lib.load().nb6_regf_device(ni, ni, di.ptr, lmax, nnbmax, 0, do.ptr, dl.ptr,
                                              None) == 0
            di.free()
            return do, dl

        do_old, dl_old = regf(rs[addr0] * rng.uniform(0.85, 1.15, ni))
        old_rows = dl_old.to_host(np.int32, (ni, lmax))
        do_old.free()
        dl_old.free()
        do, dl = regf(rs[addr0])
        force = do.to_host(np.float64, (7, ni))
        rows = dl.to_host(np.int32, (ni, lmax))
        do.free()
        dl.free()
    finally:
        g.close()
    # the caller re-runs overflowing particles with a smaller RS (util_gpu.F:65-100);
    # here they simply leave the block
    keep = (old_rows[:, 0] >= 0) & (rows[:, 0] >= 0)
    assert keep.mean() > 0.99
    addr0, old_rows, force, rows = addr0[keep], old_rows[keep], force[:, keep], rows[keep]
    ni = len(addr0)
    do = lib.DeviceArray.from_host(np.ascontiguousarray(force))
    dl = lib.DeviceArray.from_host(np.ascontiguousarray(rows))
    lst = np.zeros((n, lmax), np.int32)
    for k, i in enumerate(addr0):
        js = old_rows[k, 1:1 + old_rows[k, 0]]
        js = js[js != i] + 1
        lst[i, 0] = len(js)
        lst[i, 1:1 + len(js)] = js
    step = dtr_all / 2.0 ** rng.integers(0, 4, n)
    step[rng.random(n) < 0.2] = 2.0 ** -14
    S = dict(x=x.T.copy(), xdot=v.T.copy(), body=m, x0=(x - 0.01 * v).T.copy(),
             x0dot=v.T.copy(), f=0.5 * rng.normal(size=(3, n)), fdot=rng.normal(size=(3, n)) / 6,
             t0=time - step, h=None, step=step, timenw=time + step,
             fi=rng.normal(size=(3, n)), fidot=rng.normal(size=(3, n)),
             fr=rng.normal(size=(3, n)), frdot=rng.normal(size=(3, n)),
             d2=rng.normal(size=(3, n)), d3=rng.normal(size=(3, n)), t0r=time - dtr_all,
             stepr=dtr_all.copy(), rs=rs, list=lst,
             **{k: np.zeros((3, n)) for k in ("xn", "xndot", "d0", "d0r", "d1r", "d2r", "d3r")})
    P = dict(time=time, dtmin=2.0 ** -14, smin=2.0 ** -13, smax=0.125, etar=0.02, rscale=0.5,
             rc2=0.01, rsfac=12.0, znbmin=12.0, bodym=1.0 / n, tidal4=0.0, rdens=[0.0, 0.0, 0.0],
             n=n, ifirst=1, nnbopt=60, nnbmax=nnbmax, nb_flag=1, m_flag=0, kz38=1)
    try:
        got, cnt = _gpu(S, P, addr0 + 1, do, ni, dl, ni)
    finally:
        do.free()
        dl.free()
    sample = np.sort(rng.choice(ni, 256, replace=False))
    So = copy_state(S)
    R.regcor_block(So, P, addr0[sample] + 1, force[:, sample], rows[sample].copy())
    _same(got, So, "N=2^18 sample", idx=addr0[sample])
    # whole-block counters: NBFLUX counts every list change net of retained neighbours
    assert cnt[4] == 0 and cnt[1] > 0 and cnt[3] > 0
    flux = 0
    for k, i in enumerate(addr0):
        new = rows[k, 1:1 + rows[k, 0]]
        new = set((new[new != i] + 1).tolist())
        old = set(lst[i, 1:1 + lst[i, 0]].tolist())
        flux += len(old ^ new)
    assert cnt[1] == flux - cnt[2]
