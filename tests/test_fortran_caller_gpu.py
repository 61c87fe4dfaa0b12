"""The drop-in boundary exercised by a Fortran program (tests/fortran/regf_caller.f90,
built with amdflang by __graft_entry__.build()).  The program makes the
integrator's own calls with Fortran arrays and by-reference arguments:
GPUNB_OPEN / SEND / REGF / PROFILE / CLOSE in chunks of 1024 with
LISTGP(LMAX,1024), the overflow rescale of RS and re-run of the chunk, +IFIRST
and self removal (util_gpu.F:25-111); GPUPOT (energy.F:33); IRR_SIMD_OPEN /
SET_JP / SET_LIST / FIRR_VEC / PROFILE / CLOSE (intgrt.F:173,533,932).  Its
results are checked against the FP64 oracle, and its retry sequence against a
restatement of util_gpu.F:65-100 driven by the oracle's counts.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from helpers import vec_rel_err
from nb6 import ics

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent
EXE = HERE / "fortran" / "regf_caller"
MAXTHR = 1024
REL_TOL = 1.0e-6


def _inputs(seed=20261220, n=2048, nnbopt=50, lmax=200, nover=24):
    m, x, v = ics.plummer(n, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    rs = ics.rs_for_counts(x, np.arange(n), np.full(n, nnbopt))
    over = rng.choice(n, nover, replace=False)
    rs[over] = ics.rs_for_counts(x, over, np.full(nover, 2 * lmax))   # forced overflows
    stepr = np.minimum(0.125 * np.sqrt(1.0 + (x * x).sum(1)), 0.125) * 0.05
    ireg = np.sort(rng.choice(n, 1500, replace=False)).astype(np.int32) + 1   # 1-based
    f = 0.5 * rng.normal(0, 0.3, (n, 3))
    fdot = rng.normal(0, 0.2, (n, 3)) / 6.0
    t0 = np.floor(rng.uniform(0, 64, n)) / 1024.0
    ti = 64.0 / 1024.0
    iaddr = ireg[::2].copy()
    return dict(m=m, x=x, v=v, rs=rs, stepr=stepr, ireg=ireg, f=f, fdot=fdot, t0=t0, ti=ti,
                iaddr=iaddr, nnbopt=nnbopt, nnbmax=lmax - 50, lmax=lmax)


def _write(path, d):
    n = d["m"].size
    hdr = np.array([n, d["nnbopt"], d["nnbmax"], d["lmax"], 1, d["ireg"].size, 0, 1,
                    d["iaddr"].size], np.int32)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        for a in (d["m"], d["x"], d["v"], d["rs"], d["stepr"]):
            fh.write(np.ascontiguousarray(a, np.float64).tobytes())
        fh.write(np.float64(1.0 / n).tobytes())                    # BODYM
        for a in (d["f"], d["fdot"], d["t0"]):
            fh.write(np.ascontiguousarray(a, np.float64).tobytes())
        fh.write(np.float64(d["ti"]).tobytes())
        fh.write(d["ireg"].tobytes())
        fh.write(d["iaddr"].tobytes())


def _read(path, d):
    n, ni, na, lmax = d["m"].size, d["ireg"].size, d["iaddr"].size, d["lmax"]
    raw = Path(path).read_bytes()
    nretry = int(np.frombuffer(raw[:4], np.int32)[0])
    off = 4
    nd = 3 * ni + 3 * ni + n + n + n + 3 * na + 3 * na
    dd = np.frombuffer(raw[off:off + 8 * nd], np.float64)
    off += 8 * nd
    ints = np.frombuffer(raw[off:], np.int32)
    k = 0
    out = {}
    for name, cnt, shape in (("acc", 3 * ni, (ni, 3)), ("jrk", 3 * ni, (ni, 3)),
                             ("phidbl", n, (n,)), ("rs", n, (n,)), ("pot", n, (n,)),
                             ("ai", 3 * na, (na, 3)), ("ji", 3 * na, (na, 3))):
        out[name] = dd[k:k + cnt].reshape(shape)
        k += cnt
    out["lists"] = ints[:lmax * ni].reshape(ni, lmax)
    out["nnbid"] = ints[lmax * ni:lmax * ni + na]
    out["nretry"] = nretry
    return out


def _mirror_retries(d):
    """util_gpu.F:65-100 driven by the oracle's counts: the first overflowing row of a
    chunk rescales its RS (REAL*4 factors) and the whole chunk is re-run."""
    m, x, v = d["m"], d["x"], d["v"]
    rs = d["rs"].copy()
    nnbopt, nnbmax, lmax = d["nnbopt"], d["nnbmax"], d["lmax"]
    nretry = 0
    ni_calls = 0                     # sum of ni over the gpunb_regf calls
    for c0 in range(0, d["ireg"].size, MAXTHR):
        idx = d["ireg"][c0:c0 + MAXTHR] - 1
        while True:
            _, _, _, lst = O.regf(m, x, v, rs[idx] ** 2, d["stepr"][idx], x[idx], v[idx], lmax,
                                  nnbmax)
            ni_calls += idx.size
            bad = np.flatnonzero(lst[:, 0] < 0)
            if bad.size == 0:
                break
            k = idx[bad[0]]
            nnb = int(lst[bad[0], 0])
            if -nnb > nnbopt:
                fac = (np.float32(-np.float32(nnbopt) / np.float32(nnb))) ** np.float32(0.333)
            else:
                fac = (np.float32(nnbopt) / np.float32(nnbmax)) ** np.float32(0.4)
            rs[k] = float(np.float32(fac)) * rs[k]
            nretry += 1
    return nretry, rs, ni_calls


def test_fortran_caller_through_the_reference_interface(tmp_path):
    assert EXE.exists(), f"{EXE} is not built (__graft_entry__.build())"
    d = _inputs()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write(fin, d)
    print(f"Fortran caller: N={d['m'].size} block={d['ireg'].size} irr={d['iaddr'].size}",
          flush=True)
    p = subprocess.run([str(EXE), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    print(p.stderr[-2000:], flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    r = _read(fout, d)
    prof = _profile_lines(p.stderr)

    # the retry sequence of util_gpu.F:65-100
    nretry, rs_mirror, ni_calls = _mirror_retries(d)
    assert r["nretry"] == nretry and nretry > 0
    assert np.allclose(r["rs"], rs_mirror, rtol=4e-7, atol=0)

    # forces, potential and lists at the final radii against the FP64 oracle
    m, x, v = d["m"], d["x"], d["v"]
    idx = d["ireg"] - 1
    racc, rjrk, rpot, rlst = O.regf(m, x, v, r["rs"][idx] ** 2, d["stepr"][idx], x[idx], v[idx],
                                    d["lmax"], d["nnbmax"])
    ea, ej = vec_rel_err(r["acc"], racc), vec_rel_err(r["jrk"], rjrk)
    ep = np.abs(r["phidbl"][idx] - rpot) / rpot
    print(f"regf via Fortran: acc {ea.max():.2e} jrk {ej.max():.2e} pot {ep.max():.2e}",
          flush=True)
    assert ea.max() <= REL_TOL and ej.max() <= REL_TOL and ep.max() <= REL_TOL
    for k, i in enumerate(idx):
        c = rlst[k, 0]
        assert c >= 0
        want = [j + 1 for j in rlst[k, 1:1 + c] if j != i]       # +IFIRST, self removed
        got = r["lists"][k, 1:1 + r["lists"][k, 0]].tolist()
        assert got == want, f"row {k}"

    # GPUPOT over all N
    pp = O.pot(m, x)
    assert (np.abs(r["pot"] - pp) / pp).max() <= REL_TOL

    # irregular force on the lists the regular block produced (unlisted particles: empty)
    lists = np.zeros((m.size, d["lmax"]), np.int32)
    lists[idx] = r["lists"]
    iacc, ijrk, inn = O.firr(d["ti"], d["iaddr"], x, v, d["f"], d["fdot"], m, d["t0"], lists,
                             d["lmax"])
    ea, ej = vec_rel_err(r["ai"], iacc), vec_rel_err(r["ji"], ijrk)
    print(f"irr via Fortran: acc {ea.max():.2e} jrk {ej.max():.2e}", flush=True)
    # FP32 records and pair arithmetic of the irregular force (irr.avx.cpp's format)
    assert ea.max() <= 2e-6 and ej.max() <= 1e-5
    assert np.array_equal(r["nnbid"], inn)

    # profile lines (gpunb.velocity.cu:883-903, irr.avx.cpp:407-422): counters of the
    # calls this program made
    reg, irr = prof["reg"], prof["irr"]
    nchunks = -(-d["ireg"].size // MAXTHR)
    assert reg["Nsend"] == 1 and reg["Ngrav"] == nchunks + nretry
    assert reg["Ni"] == ni_calls                          # <Ni> = sum of ni / Nsend
    assert reg["grav"] > 0 and reg["Gflops"] > 0
    nnb = r["lists"][np.searchsorted(d["ireg"], d["iaddr"]), 0]
    assert irr["Ncall"] == 1 and irr["NI"] == int(nnb.sum())   # pairs per call
    assert abs(irr["NB"] - nnb.mean()) <= 1e-5 * nnb.mean()
    # wall time of the call (irr.avx.cpp:551,561-562), whether or not kernel timers are on;
    # usec = 1e6 grav / Ncall, both printed with six decimals
    assert irr["grav"] > 0 and irr["Gflops"] > 0
    assert abs(irr["usec"] - 1e6 * irr["grav"] / irr["Ncall"]) <= 1.0 + 1e-3 * irr["usec"]


def _profile_lines(err):
    """Parse the '[R.0 GPU Reg.F ]' and '[R.0 GPU Irr.F ]' lines."""
    import re
    out = {}
    m = re.search(r"\[R\.0 GPU Reg\.F \] Nsend (\d+)\s+Ngrav (\d+)\s+<Ni> (\d+)\s+send\(s\) (\S+) "
                  r"grav\(s\) (\S+)\s+nb\(s\) (\S+)\s+out\(s\) (\S+)\s+Perf\.\(Gflops\) (\S+)", err)
    assert m, "no gpunb_profile_ line"
    out["reg"] = dict(Nsend=int(m[1]), Ngrav=int(m[2]), Ni=int(m[3]), grav=float(m[5]),
                      Gflops=float(m[8]))
    m = re.search(r"\[R\.0 GPU Irr\.F \] Ncall: (\d+) <NI>: (\d+) <NB>: (\S+) grav: (\S+) s, "
                  r"(\S+) Gflops, (\S+) usec", err)
    assert m, "no irr_simd_profile_ line"
    out["irr"] = dict(Ncall=int(m[1]), NI=int(m[2]), NB=float(m[3]), grav=float(m[4]),
                      Gflops=float(m[5]), usec=float(m[6]))
    return out
