"""NAMELIST and unit-10 readers (nb6/io.py) on the reference's own example input
files (tests/golden/inputs/, copied from examples/) and on list-directed edge cases."""
from pathlib import Path

import numpy as np
import pytest

from nb6 import io

INP = Path(__file__).resolve().parent / "golden" / "inputs"


def test_reference_example_files_parse():
    g = io.read_namelist((INP / "N100k.inp").read_text())
    assert list(g) == ["INNBODY6", "ININPUT", "INSSE", "INBSE", "INCOLL", "INDATA", "INSETUP",
                       "INSCALE", "INXTRNL0", "INBINPOP", "INHIPOP"]
    assert g["INNBODY6"] == {"KSTART": 1, "TCOMP": 1.0e8, "TCRTP0": 1000.0, "ISERNB": 40,
                             "ISERREG": 40, "ISERKS": 0}
    i = g["ININPUT"]
    assert i["N"] == 105000 and i["NNBOPT"] == 80 and i["RBAR"] == 1.223
    assert i["DTMIN"] == 2.5e-6 and i["RMIN"] == 8.0e-5 and i["LEVEL"] == "C"
    kz = [i["KZ"][k] for k in range(1, 51)]
    assert kz[:10] == [1, 1, 1, 0, 1, 0, 3, 2, 3, 2]
    assert kz[20:30] == [1, 2, 2, 0, 2, 2, 3, 2, 0, 2]
    assert kz[40:50] == [0, 0, 0, 0, 0, 1, 3, 0, 3, 0]
    assert g["INSSE"] == {} and g["INDATA"]["NBIN0"] == 5000
    s = g["INSETUP"]
    assert s["SEMI"] is None and s["ZM1"] is None and s["RCUT"] is None
    assert s["__bare__"] == ["ZM2", "ZMH"]
    x = g["INXTRNL0"]
    assert x["GMG"] == 1.78e11 and x["RG0"] == 13.3 and x["DISK"] is None
    assert x["RG"] == {} and x["VG"] == {}
    assert g["INBINPOP"] == {"SEMI0": 0.0005, "ECC0": -1.0, "RATIO": 1.0, "RANGE": 5.0,
                             "NSKIP": 5, "IDORM": 0}


@pytest.mark.parametrize("name,n,kz22,nnbopt,tcrit", [("1k.inp", 1000, 0, 80, 30.0),
                                                     ("N1k_1Myr.inp", 1000, 0, 80, 100.0),
                                                     ("N10k_noDat10.inp", 1000, 1, 80, 1000.0),
                                                     ("N100k.inp", 105000, 2, 80, 1000.0)])
def test_run_parameters(name, n, kz22, nnbopt, tcrit):
    p = io.run_parameters(io.read_namelist((INP / name).read_text()))
    assert (p["n"], p["kz22"], p["nnbopt"], p["tcrit"]) == (n, kz22, nnbopt, tcrit)
    assert p["etai"] == 0.01 and p["etar"] == 0.01 and p["rs0"] == 0.15


def test_namelist_round_trip_and_forms():
    txt = """&G  A=1, B=2.5D-3 ! comment
    C(2:4)= 3*7, D='it''s', E=.TRUE., F=T, G2=,, H(3)=-4 /
    $OLD X=1.E2 $END"""
    g = io.read_namelist(txt)
    assert g["G"]["A"] == 1 and g["G"]["B"] == 2.5e-3
    assert g["G"]["C"] == {2: 7, 3: 7, 4: 7}
    assert g["G"]["D"] == "it's" and g["G"]["E"] is True and g["G"]["F"] is True
    assert g["G"]["G2"] == {} and g["G"]["H"] == {3: -4}
    assert g["OLD"]["X"] == 100.0
    g2 = io.read_namelist(io.write_namelist(g))
    assert g2["G"]["C"] == g["G"]["C"] and g2["G"]["B"] == g["G"]["B"]
    assert g2["G"]["D"] == "it's" and g2["OLD"] == g["OLD"]


def test_dat10_list_directed_rules():
    """Values may continue on the next line; the rest of a READ's last line is skipped;
    D exponents, repeat counts and a '/' that ends the READ (items left at zero)."""
    txt = """1.0D-3 0.1 0.2
    0.3 1.0 2.0 3.0 99 99
    2.0e-3, 2*0.5, -0.25, 3*0.0
    3e-3 1 2 3 4 /
    """
    m, x, v = io.read_dat10(txt, 3)
    assert np.array_equal(m, [1e-3, 2e-3, 3e-3])
    assert np.array_equal(x, [[0.1, 0.2, 0.3], [0.5, 0.5, -0.25], [1, 2, 3]])
    assert np.array_equal(v, [[1, 2, 3], [0, 0, 0], [4, 0, 0]])
    with pytest.raises(EOFError):
        io.read_dat10(txt, 4)


def test_dat10_round_trip(tmp_path):
    from nb6 import ics
    m, x, v = ics.plummer(500, 11)
    p = tmp_path / "dat.10"
    io.write_dat10(p, m, x, v)
    m2, x2, v2 = io.read_dat10(p)
    assert np.array_equal(m, m2) and np.array_equal(x, x2) and np.array_equal(v, v2)


def test_tree_format():
    txt = "3\n0\n0\n" + "\n".join(["0.5", "0.25", "0.25"]) + "\n" + \
          "\n".join(["1 0 0", "0 1 0", "0 0 1"]) + "\n" + "\n".join(["0 0 1", "0 1 0", "1 0 0"]) + "\n"
    m, x, v = io.read_tree(txt)
    assert np.array_equal(m, [0.5, 0.25, 0.25]) and np.array_equal(x, np.eye(3))
    assert np.array_equal(v, np.eye(3)[::-1])


def test_input_setup_from_reference_file_and_dat10(tmp_path):
    from nb6 import ics
    from nb6.rsdrive import input_setup
    m, x, v, kw, p = input_setup(INP / "1k.inp")
    assert m.size == 1000 and kw["nnbopt"] == 80 and kw["lmax"] == 400
    assert kw["dtmin"] == 2.5e-6 and kw["smax"] == 1.0 and kw["rs0"] == 0.15
    assert np.isclose(kw["eta"], 0.1)
    assert np.array_equal(m, ics.plummer(1000, p["nrand"])[0])
    # the same file with KZ(22) = 2 reads the bodies from unit 10
    text = (INP / "1k.inp").read_text().replace("KZ(21:30)=1 0 2", "KZ(21:30)=1 2 2")
    mm, xx, vv = ics.plummer(1000, 5)
    io.write_dat10(tmp_path / "dat.10", mm, xx, vv)
    m2, x2, v2, kw2, p2 = input_setup(text, tmp_path / "dat.10")
    assert p2["kz22"] == 2 and np.array_equal(x2, xx) and np.array_equal(m2, mm)
    with pytest.raises(ValueError):
        input_setup(text)
