"""ctypes binding of libnb6gpu.so (the product library) and a thin Python
mirror of the reference's host interface for this path.

The classes keep the reference's names, argument meaning and error behaviour:

* ``Gpunb``  -- gpunb_open_/send_/regf_/profile_/close_ (gpunb.velocity.cu:905-940)
* ``gpupot`` -- gpupot_ (gpupot.gpu.cu:116-127)
* ``IrrSimd`` -- irr_simd_open_/set_jp_/set_list_/firr_vec_/close_ (irr.avx.cpp:568-606)

There is no CPU fallback: if the library is missing, ``load()`` raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent.parent
LIBPATH = PKG / "libnb6gpu.so"
HEADER = PKG.parent / "include" / "nb6gpu.h"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_L = None


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _ri(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _rd(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def load():
    """Load libnb6gpu.so (raises if it has not been built)."""
    global _L
    if _L is not None:
        return _L
    if not LIBPATH.exists():
        raise RuntimeError(f"{LIBPATH} is not built; run __graft_entry__.build() or make -C "
                           f"{PKG}")
    L = ctypes.CDLL(str(LIBPATH), mode=os.RTLD_LOCAL)
    L.nb6_last_error.restype = ctypes.c_int
    L.nb6_last_error_message.restype = ctypes.c_char_p
    L.nb6_device_count.restype = ctypes.c_int
    L.nb6_kernel_ms.restype = ctypes.c_double
    L.nb6_kernel_ms.argtypes = [ctypes.c_char_p]
    L.nb6_set_kernel_timing.argtypes = [ctypes.c_int]
    L.nb6_last_regf_chunks.restype = ctypes.c_int
    L.nb6_last_regf_batches.restype = ctypes.c_int
    L.nb6_set_regf_slot_budget.argtypes = [ctypes.c_int]
    L.nb6_set_regf_refine.argtypes = [ctypes.c_int]
    L.nb6_set_regf_near.argtypes = [ctypes.c_double]
    L.nb6_set_regf_refine_weights.argtypes = [ctypes.c_double] * 4
    L.nb6_set_ens_refine_weights.argtypes = [ctypes.c_double] * 4
    L.nb6_set_regcor_pack.argtypes = [ctypes.c_int]
    L.nb6_set_regf_debug_bounds.argtypes = [_vp]
    L.nb6_last_regf_refined.restype = ctypes.c_int
    L.nb6_last_regf_refined_chunks.restype = ctypes.c_int
    L.nb6_ens_last_refined.restype = ctypes.c_int
    L.nb6_dev_alloc.restype = _vp
    L.nb6_dev_alloc.argtypes = [ctypes.c_ulonglong]
    L.nb6_dev_free.argtypes = [_vp]
    L.nb6_memcpy_htod.argtypes = [_vp, _vp, ctypes.c_ulonglong]
    L.nb6_memcpy_dtoh.argtypes = [_vp, _vp, ctypes.c_ulonglong]
    L.nb6_regf_device.argtypes = [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _vp, _vp, _vp]
    L.nb6_regf_device_noalloc.argtypes = L.nb6_regf_device.argtypes
    L.nb6_send_device.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp]
    L.nb6_pot_device.argtypes = [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
                                 _vp, _vp]
    L.nb6_set_regf_config.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.nb6_set_regf_tuning.argtypes = [ctypes.c_int, ctypes.c_int]
    L.nb6_set_regf_precision.argtypes = [ctypes.c_int]
    L.nb6_device_copy.argtypes = [_vp, _vp, ctypes.c_ulonglong, _vp]
    L.nb6_ens_send_device.argtypes = [ctypes.c_int, _ip, _vp, ctypes.c_int, _vp]
    L.nb6_ens_regf_device.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      _vp, _vp, ctypes.c_int, _vp]
    i, d, f, u = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_ulonglong
    L.nb6_irr_set_lists_device.argtypes = [i, _vp, _vp, i, i, _vp]
    L.nb6_irr_firr_device.argtypes = [d, i, _vp, _vp, _vp]
    L.nb6_set_irr_config.argtypes = [i, i]
    L.nb6_set_irr_predict.argtypes = [i]
    L.nb6_regf_autotune_device.argtypes = [i, i, _vp, i, i, i, _vp, _vp, _vp, _vp]
    L.nb6_ens_autotune.argtypes = [_vp, i, i, i, i, _vp, _vp, _vp, _vp]
    L.nb6_rsnet_open.argtypes = [u]
    L.nb6_rsnet_set_params.argtypes = [_vp]
    L.nb6_rsnet_get_params.argtypes = [_vp]
    L.nb6_rsnet_set_norm.argtypes = [_vp, _vp]
    L.nb6_rsnet_forward_device.argtypes = [i, _vp, _vp, _vp]
    L.nb6_rsnet_grad_device.argtypes = [i, _vp, _vp, _vp, _vp, _vp]
    L.nb6_rsnet_train_step_device.argtypes = [i, _vp, _vp, f, _vp, _vp]
    L.nb6_rsnet_train_epoch_device.argtypes = [i, _vp, _vp, i, f, ctypes.c_ulonglong, _vp, _vp]
    L.nb6_rs_features_device.argtypes = [i, i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.nb6_rs_predict_device.argtypes = [i, _vp, _vp, _vp, _vp, _vp, _vp, i, i, d, _vp]
    L.nb6_rs_rule_device.argtypes = [i, i, _vp, _vp, _vp, _vp, _vp, _vp, i, _vp, _vp, i, _vp, i, i,
                                     _vp, _vp]
    L.nb6_rs_kth_device.argtypes = [i, i, _vp, _vp, i, i, _vp, _vp]
    L.nb6_regcor_device.argtypes = [ctypes.POINTER(RegcorParams), ctypes.POINTER(RegState), i,
                                    _vp, _vp, i, _vp, _vp, _vp]
    L.nb6_energy_device.argtypes = [ctypes.POINTER(EnergyIn), _vp, _vp]
    _L = L
    return L


# --------------------------------------------------------------- regular corrector
NCOUNT = 7   # nb6_regcor_device counters: NBVOID NBFLUX NBSMIN NBCORR unsupported NBFAST NCMDER
# Field order of nb6_regstate / nb6_regcor_params (include/nb6gpu.h).
REG3 = ("x", "xdot", "x0", "x0dot", "f", "fdot", "fi", "fidot", "fr", "frdot", "d2", "d3",
        "xn", "xndot", "d0", "d0r", "d1r", "d2r", "d3r")      # [nmax][3] on the device
REG1 = ("body", "t0", "step", "timenw", "t0r", "stepr", "rs")  # [nmax]
_STATE_PTRS = ("x", "xdot", "body", "x0", "x0dot", "f", "fdot", "t0", "h", "step", "timenw",
               "fi", "fidot", "fr", "frdot", "d2", "d3", "t0r", "stepr", "rs", "list", "xn",
               "xndot", "d0", "d0r", "d1r", "d2r", "d3r", "listv", "r")
_PARAM_D = ("time", "dtmin", "smin", "smax", "etar", "rscale", "rc2", "rsfac", "znbmin",
            "bodym", "tidal4")
_PARAM_I = ("n", "ifirst", "nnbopt", "nnbmax", "nb_flag", "m_flag", "kz38", "kz37")


class RegState(ctypes.Structure):
    _fields_ = ([("nmax", ctypes.c_int), ("lmax", ctypes.c_int)]
                + [(k, ctypes.c_void_p) for k in _STATE_PTRS])


class RegcorParams(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_double) for k in _PARAM_D] + [("rdens", ctypes.c_double * 3)]
                + [(k, ctypes.c_int) for k in _PARAM_I] + [("cmsep2", ctypes.c_double)])


def regcor_params(P: dict) -> RegcorParams:
    p = RegcorParams()
    for k in _PARAM_D:
        setattr(p, k, float(P[k]))
    for q in range(3):
        p.rdens[q] = float(P["rdens"][q])
    for k in _PARAM_I:
        setattr(p, k, int(P.get(k, 0)))
    p.cmsep2 = float(P.get("cmsep2", 0.0))
    return p


class Regcor:
    """Device copy of the per-particle state of the regular corrector and the call
    nb6_regcor_device (REGCOR_GPU over one block, regcor_gpu.F:1-740).  S holds numpy
    arrays named as in nb6_regstate: [3][nmax] for REG3 (stored on the device as
    [nmax][3], the Fortran X(3,NMAX) layout), [nmax] for REG1, list int32
    [nmax][lmax], h [nmax - N] (optional), listv int32 and r [nmax - N] (optional,
    CHECKL with KZ(37) = 1).  Counters: int[7] (NCOUNT)."""

    def __init__(self, S: dict):
        self.L = load()
        self.nmax = int(len(S["body"]))
        self.lmax = int(S["list"].shape[1])
        self.dev = {}
        for k in REG3:
            self.dev[k] = DeviceArray.from_host(c64(np.asarray(S[k]).reshape(3, self.nmax).T))
        for k in REG1:
            self.dev[k] = DeviceArray.from_host(c64(S[k]).reshape(self.nmax))
        self.dev["list"] = DeviceArray.from_host(np.ascontiguousarray(S["list"], np.int32))
        h = S.get("h")
        self.dev["h"] = DeviceArray.from_host(c64(h)) if h is not None and len(h) else None
        lv = S.get("listv")
        self.dev["listv"] = (DeviceArray.from_host(np.ascontiguousarray(lv, np.int32))
                             if lv is not None else None)
        r = S.get("r")
        self.dev["r"] = DeviceArray.from_host(c64(r)) if r is not None and len(r) else None
        self.st = RegState(nmax=self.nmax, lmax=self.lmax)
        for k in _STATE_PTRS:
            setattr(self.st, k, self.dev[k].ptr if self.dev.get(k) is not None else None)

    def correct(self, P: dict, ni: int, d_addr, d_force, force_stride: int, d_nlist,
                d_counters, stream=None):
        p = regcor_params(P)
        rc = self.L.nb6_regcor_device(ctypes.byref(p), ctypes.byref(self.st), int(ni),
                                      d_addr.ptr, d_force.ptr, int(force_stride), d_nlist.ptr,
                                      d_counters.ptr, stream)
        if rc:
            check_error()
            raise RuntimeError(f"nb6_regcor_device returned {rc}")

    def download(self) -> dict:
        out = {k: self.dev[k].to_host(np.float64, (self.nmax, 3)).T.copy() for k in REG3}
        out.update({k: self.dev[k].to_host(np.float64, (self.nmax,)) for k in REG1})
        out["list"] = self.dev["list"].to_host(np.int32, (self.nmax, self.lmax))
        return out

    def free(self):
        for v in self.dev.values():
            if v is not None:
                v.free()


# --------------------------------------------------------------- energy
class EnergyIn(ctypes.Structure):
    _fields_ = [("nmax", ctypes.c_int), ("lmax", ctypes.c_int), ("n", ctypes.c_int),
                ("ntot", ctypes.c_int), ("npairs", ctypes.c_int),
                ("cmsep2_rmin2", ctypes.c_double)] + [
        (k, ctypes.c_void_p) for k in ("x", "xdot", "body", "list", "ht", "phidbl")]


def energy(x, xdot, body, lst, n, npairs, cmsep2_rmin2, ht=None):
    """ENERGY's GPU branch (energy.F:28-62, :119-124, phicor.f) through
    nb6_energy_device: returns (ZKIN, POT, EBIN, PHIDBL).  x, xdot [nmax][3], body
    [nmax], lst int [nmax][lmax] (count, 1-based), ht [npairs]."""
    L = load()
    nmax, lmax = len(body), lst.shape[1]
    d = {k: DeviceArray.from_host(c64(a)) for k, a in (("x", x), ("xdot", xdot), ("body", body))}
    d["list"] = DeviceArray.from_host(np.ascontiguousarray(lst, np.int32))
    d["ht"] = DeviceArray.from_host(c64(ht)) if npairs else None
    d["phidbl"] = DeviceArray(nmax * 8)
    e = EnergyIn(nmax=nmax, lmax=lmax, n=n, ntot=n + npairs, npairs=npairs,
                 cmsep2_rmin2=cmsep2_rmin2)
    for k, v in d.items():
        setattr(e, k, v.ptr if v is not None else None)
    out = (ctypes.c_double * 3)()
    try:
        rc = L.nb6_energy_device(ctypes.byref(e), ctypes.cast(out, ctypes.c_void_p), None)
        if rc:
            check_error()
            raise RuntimeError(f"nb6_energy_device returned {rc}")
        phi = d["phidbl"].to_host(np.float64, (nmax,))
    finally:
        for v in d.values():
            if v is not None:
                v.free()
    return out[0], out[1], out[2], phi


def header_symbols() -> list[str]:
    """Function names declared in include/nb6gpu.h."""
    import re
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt.split("extern \"C\" {", 1)[1])


def check_error():
    L = load()
    code = L.nb6_last_error()
    if code:
        msg = L.nb6_last_error_message().decode()
        L.nb6_clear_error()
        raise RuntimeError(f"nb6gpu error {code}: {msg}")


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Gpunb:
    """Mirror of the reference regular-force library (one instance at a time)."""

    def __init__(self, nbmax: int, rank: int = 0):
        self.L = load()
        self.rank = rank
        self.L.gpunb_open_(_ri(nbmax), _ri(rank))
        check_error()

    def send(self, m, x, v):
        m, x, v = c64(m), c64(x), c64(v)
        self.L.gpunb_send_(_ri(m.size), _d(m), _d(x), _d(v))
        check_error()

    def regf(self, h2, dtr, xi, vi, lmax, nnbmax, m_flag=0):
        h2, dtr, xi, vi = c64(h2), c64(dtr), c64(xi), c64(vi)
        ni = h2.size
        acc = np.zeros((ni, 3)); jrk = np.zeros((ni, 3)); pot = np.zeros(ni)
        lst = np.zeros((ni, lmax), dtype=np.int32)
        self.L.gpunb_regf_(_ri(ni), _d(h2), _d(dtr), _d(xi), _d(vi), _d(acc), _d(jrk), _d(pot),
                           _ri(lmax), _ri(nnbmax), _i(lst), _ri(m_flag))
        check_error()
        return acc, jrk, pot, lst

    def profile(self):
        self.L.gpunb_profile_(_ri(self.rank))

    def close(self):
        self.L.gpunb_close_()


def gpupot(m, x, istart=1, ni=None, rank=0):
    L = load()
    m, x = c64(m), c64(x)
    n = m.size
    ni = n - (istart - 1) if ni is None else ni
    out = np.zeros(max(ni, 1))
    L.gpupot_(_ri(rank), _ri(istart), _ri(ni), _ri(n), _d(m), _d(x), _d(out))
    check_error()
    return out[:ni]


class IrrSimd:
    """Mirror of the reference irregular-force library (irr_simd_* / gpuirr_*)."""

    def __init__(self, nmax, lmax, rank=0, alias=False):
        self.L = load()
        self.rank = rank
        self.alias = alias
        if alias:
            self.L.gpuirr_open_(_ri(nmax), _ri(lmax))
        else:
            self.L.irr_simd_open_(_ri(nmax), _ri(lmax), _ri(rank))
        check_error()

    def set_jp(self, addr, pos, vel, acc2, jrk6, mass, time):
        f = self.L.gpuirr_set_jp_ if self.alias else self.L.irr_simd_set_jp_
        f(_ri(addr), _d(c64(pos)), _d(c64(vel)), _d(c64(acc2)), _d(c64(jrk6)), _rd(mass), _rd(time))
        check_error()

    def set_list(self, addr, row):
        row = np.ascontiguousarray(row, dtype=np.int32)
        f = self.L.gpuirr_set_list_ if self.alias else self.L.irr_simd_set_list_
        f(_ri(addr), _i(row))
        check_error()

    def firr_vec(self, ti, addr):
        addr = np.ascontiguousarray(addr, dtype=np.int32)
        ni = addr.size
        acc = np.zeros((ni, 3)); jrk = np.zeros((ni, 3)); nn = np.zeros(ni, np.int32)
        f = self.L.gpuirr_firr_vec_ if self.alias else self.L.irr_simd_firr_vec_
        f(_rd(ti), _ri(ni), _i(addr), _d(acc), _d(jrk), _i(nn))
        check_error()
        return acc, jrk, nn

    def profile(self):
        self.L.irr_simd_profile_(_ri(self.rank))

    def close(self):
        if self.alias:
            self.L.gpuirr_close_()
        else:
            self.L.irr_simd_close_(_ri(self.rank))


class JPred:
    """Device-resident predictor (nb6_jp_*): upload corrected particles, predict all
    of them to a block time on the GPU and run gpunb_regf for blocks by address.
    Needs an open Gpunb."""

    def __init__(self, nmax: int):
        self.L = load()
        self.L.nb6_jp_open_(_ri(nmax))
        check_error()

    def set(self, addr, x0, v0, f, fdot, m, t0):
        addr = np.ascontiguousarray(addr, np.int32)
        x0, v0, f, fdot, m, t0 = (c64(a) for a in (x0, v0, f, fdot, m, t0))
        self.L.nb6_jp_set_(_ri(addr.size), _i(addr), _d(x0), _d(v0), _d(f), _d(fdot), _d(m), _d(t0))
        check_error()

    def predict_send(self, time, nj):
        self.L.nb6_jp_predict_send_(_rd(time), _ri(nj))
        check_error()

    def regf(self, addr, h2, dtr, lmax, nnbmax, m_flag=0):
        addr = np.ascontiguousarray(addr, np.int32)
        h2, dtr = c64(h2), c64(dtr)
        ni = addr.size
        acc = np.zeros((ni, 3)); jrk = np.zeros((ni, 3)); pot = np.zeros(ni)
        lst = np.zeros((ni, lmax), dtype=np.int32)
        self.L.nb6_jp_regf_(_ri(ni), _i(addr), _d(h2), _d(dtr), _d(acc), _d(jrk), _d(pot),
                            _ri(lmax), _ri(nnbmax), _i(lst), _ri(m_flag))
        check_error()
        return acc, jrk, pot, lst

    def close(self):
        self.L.nb6_jp_close_()


class DeviceArray:
    """Minimal owned device buffer for the device-resident entry points."""

    def __init__(self, nbytes: int):
        self.L = load()
        self.nbytes = int(nbytes)
        self.ptr = self.L.nb6_dev_alloc(self.nbytes)
        if not self.ptr:
            check_error()
            raise RuntimeError("device allocation failed")

    @classmethod
    def from_host(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.L.nb6_memcpy_htod(d.ptr, a.ctypes.data, a.nbytes)
        return d

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        self.L.nb6_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes)
        return out

    def free(self):
        if self.ptr:
            self.L.nb6_dev_free(self.ptr)
            self.ptr = None
