"""CPU-side checks of the C ABI library: it builds, loads, and exports every
symbol include/focalors_poisson.h declares.  No device calls."""
import ctypes
import os

from focalors_poisson_amd import _capi


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_capi.LIB_PATH), "libfocalors_poisson.so missing: run __graft_entry__.build()"
    L = ctypes.CDLL(_capi.LIB_PATH)
    declared = _capi.declared_symbols()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_status_strings_and_version():
    L = _capi.lib()
    assert L.fp_version() >= 1
    for st in range(7):
        assert _capi.status_str(st)
    assert "unknown" in _capi.status_str(99)


def test_library_carries_gfx950_code():
    data = open(_capi.LIB_PATH, "rb").read()
    assert b"gfx950" in data
