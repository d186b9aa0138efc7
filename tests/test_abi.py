"""The C-ABI library loads and exports every entry point declared in include/mmh.h (CPU-only:
no compute calls); without a GPU the context constructor fails with a status code, not a crash."""
import ctypes as C
import os
import subprocess

import pytest

import mmh_amd


def test_library_exports_header_symbols():
    assert os.path.exists(mmh_amd.LIB_PATH), "libmmh_amd.so not built"
    L = mmh_amd.lib()
    names = mmh_amd.header_symbols()
    assert len(names) >= 40
    missing = [s for s in names if not hasattr(L, s)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "-n", mmh_amd.LIB_PATH], capture_output=True, text=True)
    # the fat binary is embedded; check the offload bundle names gfx950
    data = open(mmh_amd.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_error_codes_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    c = C.c_void_p()
    rc = mmh_amd.lib().mmh_create(C.byref(c), 0)
    assert rc == mmh_amd.MMH_EDEVICE
    assert mmh_amd.lib().mmh_destroy(None) == mmh_amd.MMH_EINVAL


def test_fortran_layer_exports_reference_names():
    if not os.path.exists(mmh_amd.FLIB_PATH):
        pytest.skip("flang not present: Fortran layer not built")
    out = subprocess.run(["nm", "-D", "--defined-only", mmh_amd.FLIB_PATH], capture_output=True, text=True).stdout
    for sym in ("multigrid_", "relax_", "calcdefp_", "relaxing_", "scalar2gradient_", "defp2delta_", "nsmooth_"):
        assert (" T " + sym) in out, sym
