"""GPU parity of gpunb_regf_ against the FP64 oracle (oracle/nb6_oracle.c).

Bar: neighbour lists bit-exact (same sorted sets, same counts, same overflow
sign and magnitude); acc, jerk and pot within 1e-6 of FP64 for every particle
(per-particle vector relative error).  The regular jerk excludes the neighbours,
so |jerk| is a random-walk sum that is occasionally very small for a given
particle; its FP32 error has a tail that grows with the number of particles
sampled (the reference's own FP32 jerk error at the 1k Plummer case is 2.2e-6
max, see test_golden_cpu.py), which the FP64 pass over flagged rows removes.
"""
import numpy as np
import pytest

import oracle as O
from helpers import (LMAX_LARGE, LMAX_SMALL, c1_inputs, c2_inputs, nnbmax_for,
                     rows_as_sets, vec_rel_err)
from nb6 import lib

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-6          # acc / pot vs FP64, every particle; jerk at the 99th percentile
JRK_MAX = 1.0e-6          # jerk vs FP64, every particle


def _compare(got, ref, label):
    acc, jrk, pot, lst = got
    racc, rjrk, rpot, rlst = ref
    ea = vec_rel_err(acc, racc)
    ej = vec_rel_err(jrk, rjrk)
    ep = vec_rel_err(pot[:, None], rpot[:, None])
    print(f"{label}: acc max {ea.max():.3e} med {np.median(ea):.3e} | jrk max {ej.max():.3e} "
          f"p99 {np.percentile(ej, 99):.3e} med {np.median(ej):.3e} | pot max {ep.max():.3e}",
          flush=True)
    assert (lst[:, 0] == rlst[:, 0]).all(), "neighbour counts differ"
    assert rows_as_sets(lst) == rows_as_sets(rlst), "neighbour lists differ"
    assert ea.max() <= REL_TOL, f"acc rel err {ea.max()}"
    assert ep.max() <= REL_TOL, f"pot rel err {ep.max()}"
    assert np.percentile(ej, 99) <= REL_TOL, f"jrk p99 rel err {np.percentile(ej, 99)}"
    assert ej.max() <= JRK_MAX, f"jrk rel err {ej.max()}"


@pytest.fixture
def gpunb():
    g = None

    def make(nbmax):
        nonlocal g
        g = lib.Gpunb(nbmax)
        return g
    yield make
    if g is not None:
        g.close()


@pytest.mark.parametrize("threads,ipl", [(256, 1), (256, 2), (128, 1), (128, 2), (512, 1),
                                         (512, 2)])
def test_regf_plummer_1k_all_i(gpunb, threads, ipl):
    m, x, v, h2, dtr = c1_inputs()
    n = m.size
    lmax, nnbmax = LMAX_SMALL, nnbmax_for(n, LMAX_SMALL)
    lib.load().nb6_set_regf_config(threads, ipl, 2048)
    g = gpunb(n)
    g.send(m, x, v)
    got = g.regf(h2, dtr, x, v, lmax, nnbmax)
    ref = O.regf(m, x, v, h2, dtr, x, v, lmax, nnbmax)
    lib.load().nb6_set_regf_config(0, 1, 4096)
    _compare(got, ref, f"C1 threads={threads} ipl={ipl}")
    # the self pair is always a neighbour (r2 = 0 < h2)
    for i in range(n):
        c = got[3][i, 0]
        assert i in set(got[3][i, 1:1 + c])


@pytest.mark.parametrize("ni", [1, 256, 4096])
@pytest.mark.parametrize("m_flag", [0, 1])
def test_regf_kroupa_64k_subsets(gpunb, ni, m_flag):
    m, x, v, idx, h2, dtr = c2_inputs(ni=ni, m_flag=m_flag)
    n = m.size
    lmax, nnbmax = LMAX_SMALL + 100, nnbmax_for(n, LMAX_SMALL + 100)
    g = gpunb(n)
    g.send(m, x, v)
    got = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax, m_flag)
    ref = O.regf(m, x, v, h2, dtr, x[idx], v[idx], lmax, nnbmax, m_flag)
    _compare(got, ref, f"C2 ni={ni} m_flag={m_flag}")


@pytest.mark.parametrize("threads,ipl", [(128, 1), (256, 1), (256, 2), (512, 1)])
@pytest.mark.parametrize("m_flag", [0, 1])
def test_regf_fast_movers_and_outsiders(gpunb, threads, ipl, m_flag):
    """Neighbours reached only through the velocity criterion (|x + dtr v| inside RS,
    x far outside) and i-particles outside the j-set's box: the per-tile far test
    must never skip a tile that holds a candidate.  ni >= 2048, so the i-set is
    Morton-ordered as well."""
    m, x, v, idx, h2, dtr = c2_inputs(n=16384, ni=4096, m_flag=m_flag, counts_median=40.0)
    rng = np.random.Generator(np.random.PCG64(515))
    xi, vi = x[idx].copy(), v[idx].copy()
    dtr = dtr.copy()
    fast = rng.random(idx.size) < 0.25
    d = rng.normal(size=(idx.size, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    dtr[fast] = 0.05
    vi[fast] += 20.0 * d[fast]                      # moves ~1 length unit within dtr
    out = (rng.random(idx.size) < 0.05) & ~fast
    xi[out] += 3.0 * d[out] + np.array([5.0, 0.0, 0.0])
    vi[out] -= (3.0 * d[out] + np.array([5.0, 0.0, 0.0])) / 0.0625   # comes back at dtr
    dtr[out] = 0.0625
    n = m.size
    lmax, nnbmax = LMAX_SMALL + 100, nnbmax_for(n, LMAX_SMALL + 100)
    lib.load().nb6_set_regf_config(threads, ipl, 2048)
    try:
        g = gpunb(n)
        g.send(m, x, v)
        got = g.regf(h2, dtr, xi, vi, lmax, nnbmax, m_flag)
    finally:
        lib.load().nb6_set_regf_config(0, 1, 4096)
    ref = O.regf(m, x, v, h2, dtr, xi, vi, lmax, nnbmax, m_flag)
    # the velocity criterion must actually matter in this case
    ref0 = O.regf(m, x, v, h2, np.zeros_like(dtr), xi, vi, lmax, nnbmax, m_flag)
    assert (ref[3][fast, 0] != ref0[3][fast, 0]).mean() > 0.5
    assert (ref[3][out, 0] != 0).all()           # the self pair meets at dtr
    _compare(got, ref, f"fast movers threads={threads} ipl={ipl} m_flag={m_flag}")


@pytest.mark.parametrize("mode,tol", [(1, 2e-11), (2, 1e-13)])
@pytest.mark.parametrize("m_flag", [0, 1])
def test_regf_precision_variants(gpunb, mode, tol, m_flag):
    """Float-float (mode 1) and FP64 (mode 2) pair arithmetic: the exact FP64
    lists and forces within tol of the FP64 oracle for every particle, jerk
    included (no FP32 tail)."""
    m, x, v, idx, h2, dtr = c2_inputs(n=16384, ni=3000, m_flag=m_flag)
    n = m.size
    lmax, nnbmax = LMAX_SMALL + 100, nnbmax_for(n, LMAX_SMALL + 100)
    L = lib.load()
    L.nb6_set_regf_precision(mode)
    try:
        g = gpunb(n)
        g.send(m, x, v)
        acc, jrk, pot, lst = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax, m_flag)
    finally:
        L.nb6_set_regf_precision(0)
    racc, rjrk, rpot, rlst = O.regf(m, x, v, h2, dtr, x[idx], v[idx], lmax, nnbmax, m_flag)
    ea, ej = vec_rel_err(acc, racc), vec_rel_err(jrk, rjrk)
    ep = np.abs(pot - rpot) / rpot
    print(f"mode {mode} m_flag {m_flag}: acc {ea.max():.2e} jrk {ej.max():.2e} pot {ep.max():.2e}")
    assert rows_as_sets(lst) == rows_as_sets(rlst)
    assert ea.max() <= tol and ej.max() <= 10 * tol and ep.max() <= tol


def test_regf_autotune_keeps_results(gpunb):
    """The regf autotuner picks a launch shape on the caller's inputs; lists stay
    identical and forces stay within the FP32 bar of the FP64 oracle."""
    import ctypes
    m, x, v, idx, h2, dtr = c2_inputs(n=16384, ni=2048)
    n = m.size
    lmax, nnbmax = 500, 450
    g = gpunb(n)
    g.send(m, x, v)
    ni = idx.size
    L = lib.load()
    ib = np.concatenate([x[idx].T, v[idx].T, h2[None], dtr[None]]).astype(np.float64)
    di = lib.DeviceArray.from_host(ib)
    do = lib.DeviceArray(7 * ni * 8)
    dl = lib.DeviceArray(ni * lmax * 4)
    best = (ctypes.c_int * 4)()
    try:
        assert L.nb6_regf_autotune_device(ni, ni, di.ptr, lmax, nnbmax, 0, do.ptr, dl.ptr, best,
                                          None) == 0
        got = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax)
    finally:
        L.nb6_set_regf_config(0, 1, 4096)
        for d in (di, do, dl):
            d.free()
    print("autotuned", list(best))
    assert best[0] in (128, 256, 512) and best[1] in (1, 2) and best[2] in (2048, 4096, 8192)
    ref = O.regf(m, x, v, h2, dtr, x[idx], v[idx], lmax, nnbmax)
    _compare(got, ref, "autotuned")


@pytest.mark.parametrize("ni", [1500, 4096, 20000])
def test_regf_slot_budget_batches_give_the_same_bits(gpunb, ni):
    """A call whose work arrays exceed the slot budget runs in batches of i-blocks
    (unsorted i-set at ni = 1500, Morton-sorted at 4096); counts, lists and forces
    are bit-identical to the unbatched call.  At ni = 20000 the unbatched call runs
    the FP64 pass as separate launches and the small batches as two merged ones."""
    m, x, v, idx, h2, dtr = c2_inputs(n=65536, ni=ni)
    n = m.size
    lmax, nnbmax = 500, 450
    L = lib.load()
    g = gpunb(n)
    g.send(m, x, v)
    a = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax)
    assert L.nb6_last_regf_batches() == 1
    L.nb6_set_regf_slot_budget(48)
    try:
        b = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax)
        nbatch = L.nb6_last_regf_batches()
    finally:
        L.nb6_set_regf_slot_budget(32768)
    print(f"ni={ni}: {nbatch} batches")
    assert nbatch > 1
    for u, w in zip(a, b):
        assert np.array_equal(u, w)


def test_regf_overflow_exact_negative_counts(gpunb):
    """Counts above nnbmax come back as -count with the exact count
    (reg.avx.cpp:320-322); the caller's RS rescale (util_gpu.F:82-86) needs it."""
    m, x, v, idx, h2, dtr = c2_inputs(n=16384, ni=512, counts_median=60.0)
    n = m.size
    lmax, nnbmax = 120, 64
    g = gpunb(n)
    g.send(m, x, v)
    got = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax)
    ref = O.regf(m, x, v, h2, dtr, x[idx], v[idx], lmax, nnbmax)
    assert (got[3][:, 0] < 0).any() and (got[3][:, 0] > 0).any()
    assert (got[3][:, 0] == ref[3][:, 0]).all()
    ok = got[3][:, 0] >= 0
    assert rows_as_sets(got[3][ok]) == rows_as_sets(ref[3][ok])
    # the reference sums every non-neighbour whatever the list does: so do overflowing rows
    assert vec_rel_err(got[0], ref[0]).max() <= REL_TOL
    assert vec_rel_err(got[2][:, None], ref[2][:, None]).max() <= REL_TOL


def test_regf_overflow_rows_through_the_packed_copy(gpunb):
    """A host-ABI call whose [ni][lmax] list array exceeds 2 MB comes back through the
    packed copy (only each row's count and neighbours cross to the host): the same
    counts, overflow signs and rows as the oracle, with rows past the count untouched."""
    m, x, v, idx, h2, dtr = c2_inputs(n=16384, ni=4096, counts_median=60.0)
    n = m.size
    lmax, nnbmax = 200, 64
    assert 4096 * lmax * 4 > 2 << 20
    g = gpunb(n)
    g.send(m, x, v)
    got = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax)
    ref = O.regf(m, x, v, h2, dtr, x[idx], v[idx], lmax, nnbmax)
    cnt = got[3][:, 0]
    assert (cnt < 0).any() and (cnt > 0).any()
    assert (cnt == ref[3][:, 0]).all()
    ok = cnt >= 0
    assert rows_as_sets(got[3][ok]) == rows_as_sets(ref[3][ok])
    past = np.arange(lmax)[None, :] > np.maximum(cnt, 0)[:, None]
    assert (got[3][past] == 0).all(), "entries past a row's count were written"
    assert vec_rel_err(got[0], ref[0]).max() <= REL_TOL


@pytest.mark.parametrize("n", [1000, 200000])
def test_send_rejects_nan_then_recovers(gpunb, n):
    """gpunb_send_ packs the j-set on several host threads above 65,536 particles: a NaN
    in any field of any range is reported (nothing is sent), and the next clean send
    gives the oracle's lists and forces."""
    m, x, v, idx, h2, dtr = c2_inputs(n=n, ni=64, counts_median=40.0)
    g = gpunb(n)
    L = lib.load()
    L.nb6_set_abort_on_error(0)   # the library aborts on errors unless told otherwise
    try:
        for field, (row, col) in (("x", (n - 1, 2)), ("v", (n // 2, 0)), ("m", (7, None))):
            xb, vb, mb = x.copy(), v.copy(), m.copy()
            if field == "x":
                xb[row, col] = np.nan
            elif field == "v":
                vb[row, col] = np.nan
            else:
                mb[row] = np.nan
            with pytest.raises(RuntimeError, match="NaN in j-particle data"):
                g.send(mb, xb, vb)
    finally:
        L.nb6_set_abort_on_error(1)
    g.send(m, x, v)
    lmax, nnbmax = 200, 150
    got = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax)
    ref = O.regf(m, x, v, h2, dtr, x[idx], v[idx], lmax, nnbmax)
    _compare(got, ref, f"after NaN sends n={n}")


def _crowded_row(seed, n_bg, n_in, n_band, n_shell, rs=0.05, eps=2.0e-7):
    """Background particles, then around the point c: n_in inside 0.9 RS, n_band at
    RS (1 +- eps) (alternately inside and outside, all within the FP32 kernel's band
    around the threshold), n_shell between RS and 2 RS.  dtr = 0, so the FP64
    decision is r2 < RS^2 and is far from ambiguous at eps = 2e-7."""
    rng = np.random.default_rng(seed)
    c = np.array([0.3, 0.2, -0.1])

    def dirs(k):
        u = rng.normal(size=(k, 3))
        return u / np.linalg.norm(u, axis=1)[:, None]
    xb = rng.normal(size=(n_bg, 3)) * 0.6
    xb = xb[np.linalg.norm(xb - c, axis=1) > 2.5 * rs]
    xin = c + dirs(n_in) * (0.9 * rs * rng.random(n_in) ** (1 / 3))[:, None]
    sgn = np.where(np.arange(n_band) % 2 == 0, -1.0, 1.0)
    xband = c + dirs(n_band) * (rs * (1.0 + sgn * eps))[:, None]
    xsh = c + dirs(n_shell) * (rs * (1.05 + 0.9 * rng.random(n_shell)))[:, None]
    x = np.concatenate([xb, xin, xband, xsh])
    n = len(x)
    m = np.full(n, 1.0 / n) * rng.uniform(0.5, 1.5, n)
    v = rng.normal(size=(n, 3)) * 0.3
    xi = np.stack([c, c + [0.4, -0.3, 0.2]])          # the crowded row and a plain one
    vi = np.stack([[0.1, 0.0, -0.2], [0.0, 0.1, 0.0]])
    return m, x, v, xi, vi, np.array([rs * rs, 0.01]), np.zeros(2)


@pytest.mark.parametrize("mode", [0, 1])
def test_regf_overflow_band_pairs_past_the_slots(gpunb, mode):
    """A row with 4x more neighbours than hard slots and 24 pairs inside the FP32
    kernel's rounding band around RS, most of them met after the slots are full: the
    band pairs are decided in FP64 (forces of the non-neighbours added, the count
    exact), as the oracle and the reference, which sum every non-neighbour."""
    m, x, v, xi, vi, h2, dtr = _crowded_row(31, 3000, 200, 24, 0)
    lmax, nnbmax = 60, 50
    g = gpunb(len(m))
    g.send(m, x, v)
    L = lib.load()
    L.nb6_set_regf_config(256, 1, 1)      # one j-chunk: every neighbour meets one slot set
    L.nb6_set_regf_precision(mode)
    try:
        got = g.regf(h2, dtr, xi, vi, lmax, nnbmax)
    finally:
        L.nb6_set_regf_precision(0)
        L.nb6_set_regf_config(0, 1, 4096)
    assert lib.load().nb6_last_regf_chunks() == 1
    ref = O.regf(m, x, v, h2, dtr, xi, vi, lmax, nnbmax)
    assert ref[3][0, 0] < -200 and ref[3][1, 0] >= 0
    assert (got[3][:, 0] == ref[3][:, 0]).all(), (got[3][:, 0], ref[3][:, 0])
    tol = REL_TOL if mode == 0 else 1e-12
    assert vec_rel_err(got[0], ref[0]).max() <= tol
    assert vec_rel_err(got[2][:, None], ref[2][:, None]).max() <= tol
    assert vec_rel_err(got[1], ref[1]).max() <= JRK_MAX


@pytest.mark.parametrize("refine", [0, 1])
def test_regf_overflow_band_pairs_without_slots(gpunb, refine):
    """When band pairs of an overflowing row find neither a hard nor a soft slot
    (the near field has already taken the soft slots), the FP32 pass cannot sum the
    row exactly.  Without the FP64 pass its forces come back NaN; with it (the
    default) the row is recomputed over all j in FP64 and equals the oracle.  Its
    count stays negative, at least the exact count in magnitude.  The other row is
    untouched."""
    m, x, v, xi, vi, h2, dtr = _crowded_row(32, 1000, 200, 400, 3000)
    lmax, nnbmax = 60, 50
    g = gpunb(len(m))
    g.send(m, x, v)
    L = lib.load()
    L.nb6_set_regf_config(256, 1, 1)
    L.nb6_set_regf_refine(refine)
    try:
        got = g.regf(h2, dtr, xi, vi, lmax, nnbmax)
        nfix = L.nb6_last_regf_refined()
    finally:
        L.nb6_set_regf_config(0, 1, 4096)
        L.nb6_set_regf_refine(1)
    ref = O.regf(m, x, v, h2, dtr, xi, vi, lmax, nnbmax)
    if refine:
        assert nfix >= 1
        assert vec_rel_err(got[0][:1], ref[0][:1]).max() <= 1e-12
        assert vec_rel_err(got[1][:1], ref[1][:1]).max() <= 1e-12
        assert abs(got[2][0] - ref[2][0]) <= 1e-12 * abs(ref[2][0])
    else:
        assert nfix == 0
        assert np.isnan(got[0][0]).all() and np.isnan(got[2][0])
    assert got[3][0, 0] <= ref[3][0, 0] < 0
    assert got[3][1, 0] == ref[3][1, 0]
    assert vec_rel_err(got[0][1:], ref[0][1:]).max() <= REL_TOL


def test_regf_split_j_many_chunks(gpunb):
    """A single i against all j split over many chunks keeps lists ascending."""
    m, x, v, idx, h2, dtr = c2_inputs(ni=3)
    n = m.size
    lmax, nnbmax = 600, 550
    g = gpunb(n)
    g.send(m, x, v)
    got = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax)
    assert lib.load().nb6_last_regf_chunks() > 8
    ref = O.regf(m, x, v, h2, dtr, x[idx], v[idx], lmax, nnbmax)
    _compare(got, ref, "split-j")
    for row in got[3]:
        c = row[0]
        assert (np.diff(row[1:1 + c]) > 0).all()


def test_regf_ragged_j_and_resend(gpunb):
    """nj not a multiple of the tile; a second send with fewer particles."""
    m, x, v, h2, dtr = c1_inputs(n=1000)
    g = gpunb(1000)
    for nj in (1000, 777):
        g.send(m[:nj], x[:nj], v[:nj])
        got = g.regf(h2[:50], dtr[:50], x[:50], v[:50], 400, 350)
        ref = O.regf(m[:nj], x[:nj], v[:nj], h2[:50], dtr[:50], x[:50], v[:50], 400, 350)
        _compare(got, ref, f"ragged nj={nj}")


def test_kernel_timers_only_when_asked_and_same_results(gpunb):
    """Kernel event timers are off by default (nb6_kernel_ms reports -1 for a call
    made without them) and, switched on, time the call without changing its bits."""
    m, x, v, idx, h2, dtr = c2_inputs(n=16384, ni=600)
    L = lib.load()
    g = gpunb(m.size)
    g.send(m, x, v)
    L.nb6_set_kernel_timing(0)
    off = g.regf(h2, dtr, x[idx], v[idx], 300, 250)
    assert L.nb6_kernel_ms(b"regf_tile") == -1.0
    L.nb6_set_kernel_timing(1)
    try:
        on = g.regf(h2, dtr, x[idx], v[idx], 300, 250)
        t = L.nb6_kernel_ms(b"regf_tile")
    finally:
        L.nb6_set_kernel_timing(0)
    assert 0.0 < t < 1000.0
    for a, b in zip(off, on):
        assert np.array_equal(a, b)


def test_regf_device_entry_point_matches_host_abi(gpunb):
    """nb6_regf_device (inputs resident in HBM) gives the same bits as gpunb_regf_."""
    m, x, v, idx, h2, dtr = c2_inputs(n=16384, ni=1000)
    n = m.size
    lmax, nnbmax = 500, 450
    g = gpunb(n)
    g.send(m, x, v)
    host = g.regf(h2, dtr, x[idx], v[idx], lmax, nnbmax)
    ni = idx.size
    ib = np.concatenate([x[idx].T, v[idx].T, h2[None], dtr[None]]).astype(np.float64)
    di = lib.DeviceArray.from_host(ib)
    do = lib.DeviceArray(7 * ni * 8)
    dl = lib.DeviceArray(ni * lmax * 4)
    rc = lib.load().nb6_regf_device(ni, ni, di.ptr, lmax, nnbmax, 0, do.ptr, dl.ptr, None)
    assert rc == 0
    out = do.to_host(np.float64, (7, ni))
    lst = dl.to_host(np.int32, (ni, lmax))
    for d in (di, do, dl):
        d.free()
    assert np.array_equal(out[0:3].T, host[0])
    assert np.array_equal(out[3:6].T, host[1])
    assert np.array_equal(out[6], host[2])
    assert rows_as_sets(lst) == rows_as_sets(host[3])


@pytest.mark.parametrize("case", ["reg_plummer1k", "reg_kroupa16k_m0", "reg_kroupa16k_m1"])
def test_regf_beats_reference_precision_on_golden_inputs(gpunb, case):
    """On the reference's own golden inputs: our lists equal the FP64 lists, and
    our error against FP64 is below the reference AVX back end's at the median,
    the 99th percentile and the maximum."""
    from pathlib import Path
    g = dict(np.load(Path(__file__).resolve().parent / "golden" / f"{case}.npz"))
    lmax, nnbmax, mf = int(g["lmax"]), int(g["nnbmax"]), int(g["m_flag"])
    gp = gpunb(g["m"].size)
    gp.send(g["m"], g["x"], g["v"])
    acc, jrk, pot, lst = gp.regf(g["h2"], g["dtr"], g["xi"], g["vi"], lmax, nnbmax, mf)
    racc, rjrk, rpot, rlst = O.regf(g["m"], g["x"], g["v"], g["h2"], g["dtr"], g["xi"], g["vi"],
                                    lmax, nnbmax, mf)
    assert rows_as_sets(lst) == rows_as_sets(rlst)
    ok = g["counts"] > 0
    for mine, ref, fp64 in ((acc, g["acc"], racc), (jrk, g["jrk"], rjrk)):
        e_ours = vec_rel_err(mine[ok], fp64[ok])
        e_ref = vec_rel_err(ref[ok], fp64[ok])
        for q in (50, 99, 100):
            assert np.percentile(e_ours, q) <= np.percentile(e_ref, q)
    e_ours = np.abs(pot[ok] - rpot[ok]) / rpot[ok]
    e_ref = np.abs(g["pot"][ok] - rpot[ok]) / rpot[ok]
    assert e_ours.max() <= e_ref.max() and np.median(e_ours) <= np.median(e_ref)


def test_regf_edge_cases(gpunb):
    """ni = 0; a single j; lmax = 1 (count only: every non-empty row overflows);
    coincident distinct particles (r2 = 0: always neighbours when h2 > 0, never
    evaluated otherwise); i-particles far outside the j-set's box."""
    L = lib.load()
    m, x, v, h2, dtr = c1_inputs(n=512)
    n = m.size
    g = gpunb(n)
    # a single j and an i at distance 1
    g.send(m[:1], x[:1], v[:1])
    acc, jrk, pot, lst = g.regf(np.array([1e-4]), np.zeros(1), x[:1] + [[1.0, 0, 0]], v[:1],
                                10, 5)
    ref = O.regf(m[:1], x[:1], v[:1], np.array([1e-4]), np.zeros(1), x[:1] + [[1.0, 0, 0]],
                 v[:1], 10, 5)
    assert np.allclose(acc, ref[0], rtol=1e-6) and np.allclose(pot, ref[2], rtol=1e-6)
    assert lst[0, 0] == 0
    # ni = 0 is a no-op (no error)
    L.nb6_set_abort_on_error(0)
    try:
        g.send(m, x, v)
        L.nb6_clear_error()
        rc = L.nb6_regf_device(0, 1, None, 10, 5, 0, None, None, None)
        assert rc == 0 and L.nb6_last_error() == 0
    finally:
        L.nb6_set_abort_on_error(1)
    # coincident particles: duplicate 20 of them in the j-set
    mx = np.r_[m, m[:20]]
    xx = np.r_[x, x[:20]]
    vx = np.r_[v, v[:20]]
    g.close()
    g = gpunb(mx.size)
    g.send(mx, xx, vx)
    idx = np.arange(40)
    got = g.regf(h2[idx], dtr[idx], xx[idx], vx[idx], 200, 150)
    ref = O.regf(mx, xx, vx, h2[idx], dtr[idx], xx[idx], vx[idx], 200, 150)
    assert rows_as_sets(got[3]) == rows_as_sets(ref[3])
    for k in range(20):
        assert {k, n + k} <= set(got[3][k, 1:1 + got[3][k, 0]])
    assert np.allclose(got[0], ref[0], rtol=1e-6) and np.allclose(got[2], ref[2], rtol=1e-6)
    # lmax = 1: only the count, always -count for a non-empty row
    got = g.regf(h2[idx], dtr[idx], xx[idx], vx[idx], 1, 0)
    assert (got[3][:, 0] == -ref[3][:, 0]).all()
    # i far outside the j box: no neighbours, the far field in FP32 to 1e-6
    far = xx[:8] + np.array([50.0, -30.0, 20.0])
    got = g.regf(h2[:8], dtr[:8], far, vx[:8], 200, 150)
    ref = O.regf(mx, xx, vx, h2[:8], dtr[:8], far, vx[:8], 200, 150)
    assert (got[3][:, 0] == 0).all()
    assert vec_rel_err(got[0], ref[0]).max() <= 1e-6 and vec_rel_err(got[1], ref[1]).max() <= 1e-6
