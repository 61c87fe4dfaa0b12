"""GPU energy path (nb6_energy_device: GPUPOT + c.m. copy + PHICOR + POT/ZKIN/EBIN
sums) against the FP64 restatement of energy.F/phicor.f.  The potential kernel is
FP32 pair arithmetic on FP64 separations (relative error ~1e-7, the 1e-6 bound of the
potential); PHICOR and the sums are FP64."""
import numpy as np
import pytest

import energy as E
from helpers import energy_case
from nb6 import lib

pytestmark = pytest.mark.gpu

PHI_TOL = 1e-6


@pytest.mark.parametrize("npairs,full,thr", [(0, False, 1e-4), (40, False, 1e-4), (25, True, 1e30),
                                             (25, True, 0.0)])
def test_energy_matches_oracle(npairs, full, thr):
    x, v, m, lst, n, P, ht = energy_case(seed=40 + npairs, nsing=700, npairs=npairs, full=full)
    print(f"energy case npairs={npairs} full={full} thr={thr}", flush=True)
    zk, pot, eb, phi = lib.energy(x, v, m, lst, n, P, thr, ht)
    rz, rp, rb, rphi = E.energy(x, v, m, lst, n, P, thr, ht)
    phys = np.r_[np.arange(n), np.arange(n, n + P)]
    assert np.max(np.abs(phi[phys] - rphi[phys]) / rphi[phys]) <= PHI_TOL
    assert abs(pot - rp) / rp <= PHI_TOL
    assert abs(zk - rz) / rz <= 1e-13
    assert abs(eb - rb) <= 1e-13 * max(abs(rb), 1e-300)
    # the same inputs give the same bits (fixed reduction tree)
    zk2, pot2, eb2, phi2 = lib.energy(x, v, m, lst, n, P, thr, ht)
    assert (zk2, pot2, eb2) == (zk, pot, eb) and np.array_equal(phi2, phi)


def test_energy_phicor_given_the_gpu_potential_is_fp64_exact():
    """With the oracle fed the GPU's own c.m.-level potential, PHICOR and the sums agree
    to FP64 rounding of the sums' order."""
    x, v, m, lst, n, P, ht = energy_case(seed=77, nsing=500, npairs=30)
    thr = 1e-3
    zk, pot, eb, phi = lib.energy(x, v, m, lst, n, P, thr, ht)
    raw = lib.gpupot(m[2 * P:n + P], x[2 * P:n + P])     # GPUPOT of IFIRST..NTOT, same kernel
    rz, rp, rb, rphi = E.energy(x, v, m, lst, n, P, thr, ht, phi=raw)
    assert np.array_equal(phi[:n + P], rphi[:n + P])
    assert abs(pot - rp) <= 1e-13 * rp


def test_energy_rejects_bad_lists():
    x, v, m, lst, n, P, ht = energy_case(seed=3, nsing=100, npairs=4)
    lst = lst.copy()
    lst[2 * P + 5, 0] = lst.shape[1] + 10                # count beyond the row
    L = lib.load()
    L.nb6_set_abort_on_error(0)
    try:
        with pytest.raises(RuntimeError):
            lib.energy(x, v, m, lst, n, P, 1e-4, ht)
    finally:
        L.nb6_set_abort_on_error(1)
