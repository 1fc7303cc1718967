"""The C ABI library loads and exports every function include/dpa_hip.h
declares (no GPU needed: nothing here launches a kernel)."""
import ctypes as C
import subprocess

import pytest

import dpa_amd


def test_library_exports_every_declared_symbol():
    lib = dpa_amd.load_library()
    declared = dpa_amd.declared_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(dpa_amd.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for s in dpa_amd.declared_symbols():
        assert s in exported, s          # unmangled extern "C"
    # nothing else leaks out of the library
    assert not [s for s in exported if s.startswith("_Z") and "dpah" in s]


def test_version_and_error_string():
    lib = dpa_amd.load_library()
    assert lib.dpah_version() == 1
    assert isinstance(lib.dpah_last_error(), bytes)


def test_null_arguments_are_rejected_without_touching_the_device():
    lib = dpa_amd.load_library()
    assert lib.dpah_result_round_count(None, None) == dpa_amd.ERR_INVALID
    assert b"null" in lib.dpah_last_error()
    assert lib.dpah_result_free(None) == dpa_amd.OK
    assert lib.dpah_field_release(None) == dpa_amd.OK
    assert lib.dpah_destroy(None) == dpa_amd.OK
