"""ctypes front end of the CPU checker (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module.  It wraps

* ``liboracle.so``  -- the FP64 restatement in nb6_oracle.c (orc_regf, orc_pot,
  orc_firr), and
* ``_ref/libref_{reg,pot,irr}.so`` -- the reference's own AVX back end
  (src/Main/reg.avx.cpp, pot.avx.cpp, irr.avx.cpp) compiled by ``make ref``,
  loaded RTLD_LOCAL so that its gpunb_*/gpupot_/irr_simd_* symbols never clash
  with the product library's.

Padding rules of the reference AVX code are applied here:
reg.avx.cpp reads and writes i-arrays up to the next multiple of 4
(reg.avx.cpp:204-226,308-314); pot.avx.cpp writes 8 at a time from pot+istart-1
(pot.avx.cpp:87-137), so it is always called with istart=1 on padded buffers.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _load(path: Path):
    if not path.exists():
        raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
    return ctypes.CDLL(str(path), mode=os.RTLD_LOCAL)


_orc = None


def lib():
    global _orc
    if _orc is None:
        _orc = _load(HERE / "liboracle.so")
        _orc.orc_num_threads.restype = ctypes.c_int
        _orc.orc_is_neighbour.restype = ctypes.c_int
    return _orc


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def regf(m, x, v, h2, dtr, xi, vi, lmax, nnbmax, m_flag=0):
    """FP64 regular force; returns acc[ni,3], jrk[ni,3], pot[ni], list[ni,lmax]."""
    m, x, v = c64(m), c64(x), c64(v)
    h2, dtr, xi, vi = c64(h2), c64(dtr), c64(xi), c64(vi)
    ni = h2.size
    acc = np.zeros((ni, 3)); jrk = np.zeros((ni, 3)); pot = np.zeros(ni)
    lst = np.zeros((ni, lmax), dtype=np.int32)
    lib().orc_regf(ctypes.c_int(m.size), _d(m), _d(x), _d(v), ctypes.c_int(ni),
                   _d(h2), _d(dtr), _d(xi), _d(vi), _d(acc), _d(jrk), _d(pot),
                   ctypes.c_int(lmax), ctypes.c_int(nnbmax), _i(lst),
                   ctypes.c_int(m_flag))
    return acc, jrk, pot, lst


def is_neighbour(xj, vj, mj, xi, vi, h2, dtr, m_flag=0):
    xj, vj, xi, vi = c64(xj), c64(vj), c64(xi), c64(vi)
    return bool(lib().orc_is_neighbour(_d(xj), _d(vj), ctypes.c_double(mj), _d(xi),
                                       _d(vi), ctypes.c_double(h2),
                                       ctypes.c_double(dtr), ctypes.c_int(m_flag)))


def pot(m, x, istart=1, ni=None):
    m, x = c64(m), c64(x)
    ni = m.size - (istart - 1) if ni is None else ni
    out = np.zeros(ni)
    lib().orc_pot(ctypes.c_int(m.size), _d(m), _d(x), ctypes.c_int(istart),
                  ctypes.c_int(ni), _d(out))
    return out


def firr(ti, addr, x0, v0, a2, j6, m, t0, lists, lmax):
    """FP64 irregular force.  ``addr`` 1-based; ``lists`` [nmax, lmax] int32 with
    [:,0]=nnb and 1-based entries."""
    addr = np.ascontiguousarray(addr, dtype=np.int32)
    lists = np.ascontiguousarray(lists, dtype=np.int32)
    ni = addr.size
    acc = np.zeros((ni, 3)); jrk = np.zeros((ni, 3)); nn = np.zeros(ni, np.int32)
    lib().orc_firr(ctypes.c_double(ti), ctypes.c_int(ni), _i(addr), _d(c64(x0)),
                   _d(c64(v0)), _d(c64(a2)), _d(c64(j6)), _d(c64(m)), _d(c64(t0)),
                   _i(lists), ctypes.c_int(lmax), _d(acc), _d(jrk), _i(nn))
    return acc, jrk, nn


def num_threads():
    return lib().orc_num_threads()


# ---------------------------------------------------------------------------
# the reference's own AVX back end
# ---------------------------------------------------------------------------
_refs = {}


def ref_available() -> bool:
    return all((HERE / "_ref" / f"libref_{k}.so").exists() for k in ("reg", "pot", "irr"))


def _ref(kind):
    if kind not in _refs:
        _refs[kind] = _load(HERE / "_ref" / f"libref_{kind}.so")
    return _refs[kind]


class RefReg:
    """The reference's reg.avx.cpp behind its own gpunb_* ABI."""

    def __init__(self, nbmax, rank=0):
        self.L = _ref("reg")
        self.nbmax = nbmax
        self.L.gpunb_open_(ctypes.byref(ctypes.c_int(nbmax + 4)), ctypes.byref(ctypes.c_int(rank)))

    def send(self, m, x, v):
        m, x, v = c64(m), c64(x), c64(v)
        self._keep = (m, x, v)
        self.L.gpunb_send_(ctypes.byref(ctypes.c_int(m.size)), _d(m), _d(x), _d(v))

    def regf(self, h2, dtr, xi, vi, lmax, nnbmax, m_flag=0):
        ni = len(h2)
        np4 = (ni + 3) // 4 * 4
        def pad(a, shape):
            out = np.zeros(shape)
            out[:ni] = a
            return out
        h2p = pad(h2, np4); dtp = pad(dtr, np4)
        xip = pad(xi, (np4, 3)); vip = pad(vi, (np4, 3))
        acc = np.zeros((np4, 3)); jrk = np.zeros((np4, 3)); pt = np.zeros(np4)
        lst = np.zeros((np4, lmax), dtype=np.int32)
        self.L.gpunb_regf_(ctypes.byref(ctypes.c_int(ni)), _d(h2p), _d(dtp), _d(xip),
                           _d(vip), _d(acc), _d(jrk), _d(pt),
                           ctypes.byref(ctypes.c_int(lmax)),
                           ctypes.byref(ctypes.c_int(nnbmax)), _i(lst),
                           ctypes.byref(ctypes.c_int(m_flag)))
        return acc[:ni], jrk[:ni], pt[:ni], lst[:ni]

    def close(self):
        self.L.gpunb_close_()


def ref_pot(m, x, rank=0):
    """Reference pot.avx.cpp over all particles (istart=1, ni=n)."""
    L = _ref("pot")
    m, x = c64(m), c64(x)
    n = m.size
    n8 = (n + 7) // 8 * 8 + 8
    mp = np.zeros(n8); xp = np.zeros((n8, 3))
    mp[:n] = m; xp[:n] = x
    out = np.zeros(n8)
    L.gpupot_(ctypes.byref(ctypes.c_int(rank)), ctypes.byref(ctypes.c_int(1)),
              ctypes.byref(ctypes.c_int(n)), ctypes.byref(ctypes.c_int(n)),
              _d(mp), _d(xp), _d(out))
    return out[:n]


class RefIrr:
    """The reference's irr.avx.cpp behind its irr_simd_* ABI."""

    def __init__(self, nmax, lmax, rank=0):
        self.L = _ref("irr")
        self.rank = rank
        self.L.irr_simd_open_(ctypes.byref(ctypes.c_int(nmax)), ctypes.byref(ctypes.c_int(lmax)),
                              ctypes.byref(ctypes.c_int(rank)))

    def set_jp(self, addr, pos, vel, acc2, jrk6, mass, time):
        self.L.irr_simd_set_jp_(ctypes.byref(ctypes.c_int(addr)), _d(c64(pos)), _d(c64(vel)),
                                _d(c64(acc2)), _d(c64(jrk6)),
                                ctypes.byref(ctypes.c_double(mass)),
                                ctypes.byref(ctypes.c_double(time)))

    def set_list(self, addr, row):
        # the reference reads list entries 8 at a time (irr.avx.cpp:468-489)
        nnb = int(row[0])
        buf = np.zeros(1 + (nnb + 7) // 8 * 8 + 8, dtype=np.int32)
        buf[:1 + nnb] = row[:1 + nnb]
        self.L.irr_simd_set_list_(ctypes.byref(ctypes.c_int(addr)), _i(buf))

    def firr_vec(self, ti, addr):
        addr = np.ascontiguousarray(addr, dtype=np.int32)
        ni = addr.size
        acc = np.zeros((ni, 3)); jrk = np.zeros((ni, 3)); nn = np.zeros(ni, np.int32)
        self.L.irr_simd_firr_vec_(ctypes.byref(ctypes.c_double(ti)), ctypes.byref(ctypes.c_int(ni)),
                                  _i(addr), _d(acc), _d(jrk), _i(nn))
        return acc, jrk, nn

    def close(self):
        self.L.irr_simd_close_(ctypes.byref(ctypes.c_int(self.rank)))
