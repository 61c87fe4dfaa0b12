"""Pins the FP64 restatement of regint.f (oracle/regint.py): against the
gpunb_regf oracle (same pair terms, neighbours excluded), the all-pairs potential,
and the overflow retry including its quirks (POT added once per attempt; the
NB_FLAG = 1 rescale grows RS because of Fortran precedence)."""
import numpy as np
import pytest

import oracle as O
import regint as RI
from nb6 import ics


def setup(n=600, seed=3):
    m, x, v = ics.plummer(n, seed)
    idx = np.arange(0, n, 13)
    rs = np.zeros(n)
    rs[idx] = ics.rs_for_counts(x, idx, np.full(idx.size, 25))
    dtr = np.full(n, 0.01)
    return m, x, v, idx, rs, dtr


def test_regint_matches_regf_terms_and_full_potential():
    m, x, v, idx, rs, dtr = setup()
    freg, fdr, lists, pot, rs2, retries = RI.regint(x, v, m, rs, dtr, idx, 100, 50)
    acc, jrk, _, lst = O.regf(m, x, v, rs[idx] ** 2, dtr[idx], x[idx], v[idx], 100, 80)
    assert np.allclose(freg, acc, rtol=1e-12, atol=1e-14)
    assert np.allclose(fdr, jrk, rtol=1e-11, atol=1e-13)
    full = O.pot(m, x)
    assert np.allclose(pot[idx], full[idx], rtol=1e-12)
    for k, i in enumerate(idx):     # regf lists self, regint does not
        mine = set(lists[k, 1:1 + lists[k, 0]].tolist())
        assert mine == set(lst[k, 1:1 + lst[k, 0]].tolist()) - {i}
    assert not retries.any()


def test_regint_accumulates_potential():
    m, x, v, idx, rs, dtr = setup()
    _, _, _, pot, _, _ = RI.regint(x, v, m, rs, dtr, idx, 100, 50, pot=np.full(m.size, 2.5))
    assert np.allclose(pot[idx], 2.5 + O.pot(m, x)[idx], rtol=1e-12)


def test_regint_overflow_retry_shrinks_with_nb_flag_0():
    """NB_FLAG = 0: RS x 0.9 until the count fits; POT gains one full sum per attempt."""
    m, x, v, idx, rs, dtr = setup()
    i = idx[:1]
    big = rs.copy()
    big[i] *= 4.0
    freg, fdr, lists, pot, rs2, retries = RI.regint(x, v, m, big, dtr, i, 40, 20, nb_flag=0)
    assert retries[0] > 0
    assert np.isclose(rs2[i[0]], big[i[0]] * 0.9 ** retries[0])
    assert lists[0, 0] + 1 <= 40 - 3
    assert np.isclose(pot[i[0]], (retries[0] + 1) * O.pot(m, x)[i[0]], rtol=1e-12)


def test_regint_nb_flag_1_rescale_grows_the_radius():
    """NNBOPT / NNB**0.3333 > 1 for NNB < NNBOPT**3: the retry cannot converge."""
    m, x, v, idx, rs, dtr = setup()
    i = idx[:1]
    big = rs.copy()
    big[i] *= 4.0
    with pytest.raises(RI.NoConvergence):
        RI.regint(x, v, m, big, dtr, i, 40, 20, nb_flag=1, max_retry=5)
