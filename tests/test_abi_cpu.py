"""The C-ABI library loads without a GPU and exports every symbol that
include/nb6gpu.h declares; calls that need a device fail with a status instead
of crashing when abort-on-error is switched off."""
import ctypes
import subprocess

import pytest

from nb6 import lib


def test_library_exports_every_declared_symbol():
    syms = lib.header_symbols()
    assert "gpunb_regf_" in syms and "gpupot_" in syms and "irr_simd_firr_vec_" in syms
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib.LIBPATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing


def test_reference_abi_names_present():
    """Part 1 of the header is exactly the reference's Fortran-callable set
    (gpunb.velocity.cu:905-940, gpupot.gpu.cu:116-127, irr.avx.cpp:568-606)."""
    want = {"gpunb_devinit_", "gpunb_open_", "gpunb_close_", "gpunb_send_", "gpunb_regf_",
            "gpunb_profile_", "gpupot_", "irr_simd_open_", "irr_simd_close_",
            "irr_simd_profile_", "irr_simd_set_jp_", "irr_simd_set_list_", "irr_simd_firr_vec_",
            "gpuirr_open_", "gpuirr_close_", "gpuirr_set_jp_", "gpuirr_set_list_",
            "gpuirr_firr_vec_"}
    assert want <= set(lib.header_symbols())


def test_no_device_is_an_error_status_not_a_crash():
    L = lib.load()
    if L.nb6_device_count() > 0:
        pytest.skip("a GPU is visible")
    L.nb6_set_abort_on_error(0)
    L.nb6_clear_error()
    L.gpunb_open_(ctypes.byref(ctypes.c_int(100)), ctypes.byref(ctypes.c_int(0)))
    assert L.nb6_last_error() == 4
    assert b"device" in L.nb6_last_error_message()
    L.nb6_clear_error()
    # calls in the wrong state report E_STATE
    L.gpunb_regf_(*(ctypes.byref(ctypes.c_int(1)),) + (None,) * 7 +
                  (ctypes.byref(ctypes.c_int(10)), ctypes.byref(ctypes.c_int(5)), None,
                   ctypes.byref(ctypes.c_int(0))))
    assert L.nb6_last_error() == 2
    L.nb6_clear_error()
    L.nb6_set_abort_on_error(1)
