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


def _c_layout(tmp_path, struct, fields):
    """sizeof and offsetof of a header struct, from a program gcc compiles here."""
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "nb6gpu.h"\nint main(void){\n'
                   f'printf("%zu\\n", sizeof({struct}));\n' +
                   "".join(f'printf("%zu\\n", offsetof({struct}, {f}));\n' for f in fields) +
                   "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(lib.HEADER.parent), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    return vals[0], vals[1:]


@pytest.mark.parametrize("cls,name", [(lib.RegState, "nb6_regstate"),
                                      (lib.RegcorParams, "nb6_regcor_params"),
                                      (lib.EnergyIn, "nb6_energy_in")])
def test_binding_structs_match_the_header(tmp_path, cls, name):
    fields = [f[0] for f in cls._fields_]
    size, offs = _c_layout(tmp_path, name, fields)
    assert ctypes.sizeof(cls) == size
    assert [getattr(cls, f).offset for f in fields] == offs


def test_regcor_and_energy_argument_errors():
    """Bad sizes are reported before any device work (E_ARG = 1)."""
    L = lib.load()
    L.nb6_set_abort_on_error(0)
    L.nb6_clear_error()
    try:
        p = lib.RegcorParams(time=1.0, n=10, ifirst=1, nnbmax=50, kz38=1)
        st = lib.RegState(nmax=10, lmax=40)          # lmax <= nnbmax + 1
        rc = L.nb6_regcor_device(ctypes.byref(p), ctypes.byref(st), 5, None, None, 5, None,
                                 None, None)
        assert rc == 1 and b"lmax" in L.nb6_last_error_message()
        L.nb6_clear_error()
        st = lib.RegState(nmax=10, lmax=60)
        p.kz38 = 4
        rc = L.nb6_regcor_device(ctypes.byref(p), ctypes.byref(st), 5, None, None, 5, None,
                                 None, None)
        assert rc == 1 and b"KZ(38)" in L.nb6_last_error_message()
        L.nb6_clear_error()
        assert L.nb6_regcor_device(ctypes.byref(p), ctypes.byref(st), 0, None, None, 0, None,
                                   None, None) == 0          # an empty block is a no-op
        e = lib.EnergyIn(nmax=10, lmax=4, n=10, ntot=12, npairs=1)   # NTOT != N + NPAIRS
        out = (ctypes.c_double * 3)()
        rc = L.nb6_energy_device(ctypes.byref(e), ctypes.cast(out, ctypes.c_void_p), None)
        assert rc == 1 and b"NPAIRS" in L.nb6_last_error_message()
        L.nb6_clear_error()
    finally:
        L.nb6_set_abort_on_error(1)
