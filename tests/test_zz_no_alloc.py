"""nb6_regf_device_noalloc (the regular force with no allocation: work memory in a
fixed device array) against nb6_regf_device through the C ABI: two calls in a
row, the second with a smaller i-set, each with the lists identical and the
forces within the FP32 bar of the allocating entry point's (both are within 1e-6
of FP64; they may split j into different chunks and so round differently) and
within 1e-6 of the FP64 kernel."""
import numpy as np
import pytest

from helpers import c2_inputs
from nb6 import lib

pytestmark = pytest.mark.gpu


def _same_rows(a, b):
    """Counts equal and, where non-negative, the neighbours they cover."""
    if not np.array_equal(a[:, 0], b[:, 0]):
        return False
    c = np.maximum(a[:, 0], 0)
    mask = np.arange(a.shape[1] - 1)[None, :] < c[:, None]
    return bool(np.array_equal(np.where(mask, a[:, 1:], 0), np.where(mask, b[:, 1:], 0)))


def _rel(a, b):
    return np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)


def _call(L, fn, ib, ni, lmax, nnbmax, prec=0):
    di = lib.DeviceArray.from_host(ib)
    do = lib.DeviceArray(7 * ni * 8)
    dl = lib.DeviceArray(ni * lmax * 4)
    L.nb6_set_regf_precision(prec)
    try:
        assert fn(ni, ni, di.ptr, lmax, nnbmax, 0, do.ptr, dl.ptr, None) == 0
        L.nb6_sync()
        return do.to_host(np.float64, (7, ni)), dl.to_host(np.int32, (ni, lmax))
    finally:
        L.nb6_set_regf_precision(0)
        for d in (di, do, dl):
            d.free()


def test_noalloc_equals_the_allocating_entry_point():
    L = lib.load()
    m, x, v, idx, h2, dtr = c2_inputs(n=65536, ni=5000)
    lmax, nnbmax = 500, 450
    g = lib.Gpunb(m.size)
    try:
        g.send(m, x, v)
        for ni in (5000, 700):          # the second call smaller: stale contents must not leak
            sel = slice(0, ni)
            ib = np.ascontiguousarray(np.concatenate([x[idx[sel]].T, v[idx[sel]].T,
                                                      h2[sel][None], dtr[sel][None]]))
            na_out, na_lst = _call(L, L.nb6_regf_device_noalloc, ib, ni, lmax, nnbmax)
            out, lst = _call(L, L.nb6_regf_device, ib, ni, lmax, nnbmax)
            ref, rlst = _call(L, L.nb6_regf_device, ib, ni, lmax, nnbmax, prec=2)
            ea, ej, ep = _rel(na_out[0:3], out[0:3]), _rel(na_out[3:6], out[3:6]), \
                _rel(na_out[6:7], out[6:7])
            fa, fj, fpot = _rel(na_out[0:3], ref[0:3]), _rel(na_out[3:6], ref[3:6]), \
                _rel(na_out[6:7], ref[6:7])
            print(f"noalloc ni={ni}: vs nb6_regf_device acc {ea.max():.2e} jrk {ej.max():.2e} "
                  f"pot {ep.max():.2e}; vs FP64 acc {fa.max():.2e} jrk {fj.max():.2e} "
                  f"pot {fpot.max():.2e}", flush=True)
            assert _same_rows(na_lst, lst), f"ni={ni}: lists differ"
            assert _same_rows(na_lst, rlst)
            assert max(ea.max(), ej.max(), ep.max()) <= 2e-6
            assert max(fa.max(), fj.max(), fpot.max()) <= 1e-6
    finally:
        g.close()


def test_noalloc_rejects_large_i_sets():
    L = lib.load()
    L.nb6_set_abort_on_error(0)
    g = lib.Gpunb(1024)
    try:
        m, x, v, idx, h2, dtr = c2_inputs(n=1024, ni=16)
        g.send(m, x, v)
        rc = L.nb6_regf_device_noalloc(65537, 65537, None, 100, 50, 0, None, None, None)
        assert rc != 0
        L.nb6_clear_error()
    finally:
        g.close()
        L.nb6_set_abort_on_error(1)
