"""GPU parity of gpupot_ against the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
from nb6 import ics, lib

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-6


@pytest.mark.parametrize("n", [1000, 16384])
def test_pot_plummer(n):
    m, x, v = ics.plummer(n, 7 + n)
    got = lib.gpupot(m, x)
    ref = O.pot(m, x)
    err = np.abs(got - ref) / ref
    print(f"pot n={n}: max rel err {err.max():.3e}", flush=True)
    assert err.max() <= REL_TOL


def test_pot_istart_subrange_and_binaries():
    """istart > 1 (gpupot.gpu.cu:35 indexes ptcl[i+istart-1]) and tight binaries
    whose separation is below FP32 resolution at |x| ~ 1."""
    m, x, v = ics.plummer(8192, 11, masses=ics.kroupa_masses(8192, 12))
    m, x, v = ics.add_binaries(m, x, v, 0.5, 13)
    got = lib.gpupot(m, x, istart=1001, ni=3000)
    ref = O.pot(m, x, istart=1001, ni=3000)
    err = np.abs(got - ref) / ref
    print(f"pot binaries: max rel err {err.max():.3e}", flush=True)
    assert err.max() <= REL_TOL


def test_pot_coincident_particles_skipped():
    m = np.array([0.25, 0.25, 0.5])
    x = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    got = lib.gpupot(m, x)
    assert np.allclose(got, [0.5, 0.5, 0.5], rtol=1e-7)


def test_pot_golden_binaries_beats_reference():
    """Reference AVX potential on a binary-rich cluster (tests/golden): our error
    against FP64 is below the reference's own."""
    from pathlib import Path
    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "pot_binaries4k.npz"))
    got = lib.gpupot(g["m"], g["x"])
    ref = O.pot(g["m"], g["x"])
    e_ours = np.abs(got - ref) / ref
    e_avx = np.abs(g["pot"] - ref) / ref
    assert e_ours.max() <= REL_TOL and e_ours.max() < e_avx.max()
