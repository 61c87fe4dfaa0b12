"""Batched ensemble: every cluster's result equals that cluster computed alone
(FP64 oracle parity, same bar as gpunb_regf_), in one persistent launch, and
hipGraph replays reproduce the eager step bit for bit."""
import numpy as np
import pytest

import oracle as O
from helpers import rows_as_sets, vec_rel_err
from nb6 import ics, lib

pytestmark = pytest.mark.gpu


def make_ensemble(sizes, seed=77, counts=40):
    ms, xs, vs, h2s, dts = [], [], [], [], []
    for c, n in enumerate(sizes):
        m, x, v = ics.plummer(n, seed + c, masses=ics.kroupa_masses(n, seed + 100 + c))
        rs = ics.rs_for_counts(x, np.arange(n), np.full(n, min(counts, n - 1)))
        r2 = (x * x).sum(1)
        ms.append(m); xs.append(x); vs.append(v); h2s.append(rs * rs)
        dts.append(np.minimum(0.125 * np.sqrt(1 + r2), 0.125) * 0.05)
    off = np.r_[0, np.cumsum(sizes)].astype(np.int32)
    return ms, xs, vs, h2s, dts, off


def run_ensemble(ms, xs, vs, h2s, dts, off, lmax, nnbmax, use_graph, reps=1):
    L = lib.load()
    ntot = int(off[-1])
    m, x, v = np.concatenate(ms), np.concatenate(xs), np.concatenate(vs)
    h2, dtr = np.concatenate(h2s), np.concatenate(dts)
    dj = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, v.T, m[None]])))
    di = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, v.T, h2[None], dtr[None]])))
    do = lib.DeviceArray(7 * ntot * 8)
    dl = lib.DeviceArray(ntot * lmax * 4)
    assert L.nb6_ens_open(len(off) - 1, ntot) == 0
    assert L.nb6_ens_send_device(len(off) - 1, off.ctypes.data_as(lib._ip), dj.ptr, ntot, None) == 0
    outs = []
    for _ in range(reps):
        assert L.nb6_ens_regf_device(di.ptr, ntot, lmax, nnbmax, 0, do.ptr, dl.ptr, use_graph, None) == 0
        L.nb6_sync()
        outs.append((do.to_host(np.float64, (7, ntot)), dl.to_host(np.int32, (ntot, lmax))))
    L.nb6_ens_close()
    for d in (dj, di, do, dl):
        d.free()
    return outs


def test_ensemble_matches_each_cluster_alone():
    sizes = [1000, 257, 3000, 64, 2048, 700]
    ms, xs, vs, h2s, dts, off = make_ensemble(sizes)
    lmax, nnbmax = 200, 150
    (out, lst), = run_ensemble(ms, xs, vs, h2s, dts, off, lmax, nnbmax, use_graph=0)
    for c in range(len(sizes)):
        a, b = off[c], off[c + 1]
        racc, rjrk, rpot, rlst = O.regf(ms[c], xs[c], vs[c], h2s[c], dts[c], xs[c], vs[c], lmax, nnbmax)
        assert rows_as_sets(lst[a:b]) == rows_as_sets(rlst), f"cluster {c}"
        ea = vec_rel_err(out[0:3, a:b].T, racc)
        ej = vec_rel_err(out[3:6, a:b].T, rjrk)
        ep = np.abs(out[6, a:b] - rpot) / rpot
        print(f"cluster {c} n={sizes[c]}: acc {ea.max():.2e} jrk p99 {np.percentile(ej, 99):.2e} "
              f"pot {ep.max():.2e}", flush=True)
        assert ea.max() <= 1e-6 and ep.max() <= 1e-6 and np.percentile(ej, 99) <= 1e-6


def test_ensemble_graph_replays_equal_eager():
    sizes = [512, 1500, 900]
    ms, xs, vs, h2s, dts, off = make_ensemble(sizes, seed=91)
    eager = run_ensemble(ms, xs, vs, h2s, dts, off, 120, 100, use_graph=0)[0]
    graph = run_ensemble(ms, xs, vs, h2s, dts, off, 120, 100, use_graph=1, reps=4)
    for out, lst in graph:
        assert np.array_equal(out, eager[0])
        assert rows_as_sets(lst) == rows_as_sets(eager[1])


def test_ensemble_autotune_keeps_results():
    """The autotuner only changes the launch shape: lists stay identical and the
    forces agree with the default configuration to FP32 summation-order level."""
    import ctypes
    sizes = [700, 2500, 300, 1200]
    ms, xs, vs, h2s, dts, off = make_ensemble(sizes, seed=5)
    base = run_ensemble(ms, xs, vs, h2s, dts, off, 150, 120, use_graph=0)[0]
    L = lib.load()
    ntot = int(off[-1])
    m, x, v = np.concatenate(ms), np.concatenate(xs), np.concatenate(vs)
    h2, dtr = np.concatenate(h2s), np.concatenate(dts)
    dj = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, v.T, m[None]])))
    di = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, v.T, h2[None], dtr[None]])))
    do = lib.DeviceArray(7 * ntot * 8)
    dl = lib.DeviceArray(ntot * 150 * 4)
    best = (ctypes.c_int * 4)()
    try:
        assert L.nb6_ens_open(len(sizes), ntot) == 0
        assert L.nb6_ens_send_device(len(sizes), off.ctypes.data_as(lib._ip), dj.ptr, ntot, None) == 0
        assert L.nb6_ens_autotune(di.ptr, ntot, 150, 120, 0, do.ptr, dl.ptr, best, None) == 0
        assert best[0] in (128, 256, 512) and best[1] in (1, 2) and best[2] >= 1
        assert L.nb6_ens_regf_device(di.ptr, ntot, 150, 120, 0, do.ptr, dl.ptr, 1, None) == 0
        L.nb6_sync()
        out = do.to_host(np.float64, (7, ntot))
        lst = dl.to_host(np.int32, (ntot, 150))
    finally:
        L.nb6_ens_close()
        L.nb6_set_regf_config(0, 1, 4096)
        for d in (dj, di, do, dl):
            d.free()
    print("autotuned", list(best))
    assert rows_as_sets(lst) == rows_as_sets(base[1])
    ea = vec_rel_err(out[0:3].T, base[0][0:3].T)
    assert ea.max() < 1e-6


def _same_rows(a, b):
    """Counts equal and, where non-negative, the neighbours they cover (entries past
    a row's count are not written)."""
    if not np.array_equal(a[:, 0], b[:, 0]):
        return False
    c = np.maximum(a[:, 0], 0)
    mask = np.arange(a.shape[1] - 1)[None, :] < c[:, None]
    return bool(np.array_equal(np.where(mask, a[:, 1:], 0), np.where(mask, b[:, 1:], 0)))


def test_ensemble_resend_same_offsets_equals_a_fresh_send():
    """A second send with the same offsets keeps the cluster layout and work items
    (an integration re-sends every step): the graph-replayed step on the moved
    particles gives the same bits as a freshly opened ensemble sent those positions;
    a send with different offsets rebuilds the layout."""
    sizes = [1500, 300, 5000, 64, 2600]
    ms, xs, vs, h2s, dts, off = make_ensemble(sizes, seed=91)
    lmax, nnbmax = 200, 150
    # the same clusters a step later (positions drifted along their velocities)
    xs2 = [x + 0.01 * v for x, v in zip(xs, vs)]
    L = lib.load()
    ntot = int(off[-1])

    def arrays(xl):
        m, x, v = np.concatenate(ms), np.concatenate(xl), np.concatenate(vs)
        h2, dtr = np.concatenate(h2s), np.concatenate(dts)
        dj = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, v.T, m[None]])))
        di = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, v.T, h2[None], dtr[None]])))
        return dj, di

    dj1, di1 = arrays(xs)
    dj2, di2 = arrays(xs2)
    do = lib.DeviceArray(7 * ntot * 8)
    dl = lib.DeviceArray(ntot * lmax * 4)
    offp = off.ctypes.data_as(lib._ip)
    try:
        # fresh ensemble, second positions
        assert L.nb6_ens_open(len(sizes), ntot) == 0
        assert L.nb6_ens_send_device(len(sizes), offp, dj2.ptr, ntot, None) == 0
        assert L.nb6_ens_regf_device(di2.ptr, ntot, lmax, nnbmax, 0, do.ptr, dl.ptr, 1, None) == 0
        L.nb6_sync()
        want = (do.to_host(np.float64, (7, ntot)), dl.to_host(np.int32, (ntot, lmax)))
        L.nb6_ens_close()
        # first positions, then the second with the same offsets on the kept layout
        assert L.nb6_ens_open(len(sizes), ntot) == 0
        assert L.nb6_ens_send_device(len(sizes), offp, dj1.ptr, ntot, None) == 0
        assert L.nb6_ens_regf_device(di1.ptr, ntot, lmax, nnbmax, 0, do.ptr, dl.ptr, 1, None) == 0
        assert L.nb6_ens_send_device(len(sizes), offp, dj2.ptr, ntot, None) == 0
        assert L.nb6_ens_regf_device(di2.ptr, ntot, lmax, nnbmax, 0, do.ptr, dl.ptr, 1, None) == 0
        L.nb6_sync()
        got = (do.to_host(np.float64, (7, ntot)), dl.to_host(np.int32, (ntot, lmax)))
        assert np.array_equal(got[0], want[0])
        assert _same_rows(got[1], want[1])
        # different offsets (two clusters merged): the layout is rebuilt, results stay
        # those of a direct computation on each cluster
        off3 = np.array([0, sizes[0] + sizes[1], ntot - sizes[4] - sizes[3], ntot - sizes[4], ntot],
                        dtype=np.int32)
        assert L.nb6_ens_send_device(4, off3.ctypes.data_as(lib._ip), dj2.ptr, ntot, None) == 0
        assert L.nb6_ens_regf_device(di2.ptr, ntot, lmax, nnbmax, 0, do.ptr, dl.ptr, 1, None) == 0
        L.nb6_sync()
        lst3 = dl.to_host(np.int32, (ntot, lmax))
        a, b = int(off3[3]), int(off3[4])    # the last cluster is unchanged
        assert _same_rows(lst3[a:b], want[1][a:b])
    finally:
        L.nb6_ens_close()
        for d in (dj1, di1, dj2, di2, do, dl):
            d.free()
