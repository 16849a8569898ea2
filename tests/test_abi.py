"""The C-ABI library loads and exports every entry point include/amoeba_nb.h
declares (no device calls)."""
import ctypes
import os
import re

from forcefieldx_amd import ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "amoeba_nb.h")).read()
    return sorted(set(re.findall(r"\b(nb_[a-z_]+)\s*\(", text)))


def test_library_exports_all_declared_symbols():
    L = ffi.load()
    names = declared_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(L, name), name
        assert isinstance(getattr(L, name), ctypes._CFuncPtr)


def test_abi_version_and_struct_sizes():
    L = ffi.load()
    assert L.nb_abi_version() == 9
    # the ctypes mirrors have the sizes the C compiler gives the header structs
    sizes = (ctypes.c_int32 * 3)()
    assert L.nb_struct_sizes(sizes) == ffi.NB_OK
    assert list(sizes) == [ctypes.sizeof(ffi.NbParams), ctypes.sizeof(ffi.NbTopology), ctypes.sizeof(ffi.NbEnergies)]
    assert L.nb_valence_struct_size() == ctypes.sizeof(ffi.NbValence)


def test_invalid_arguments_are_rejected_without_device_work():
    L = ffi.load()
    h = ctypes.c_void_p()
    assert L.nb_create(None, None, ctypes.byref(h)) == ffi.NB_ERR_INVALID
    assert L.nb_set_cell(None, None) == ffi.NB_ERR_INVALID
    assert L.nb_evaluate(None, 0, None) == ffi.NB_ERR_INVALID
    assert L.nb_set_valence(None, None) == ffi.NB_ERR_INVALID
    assert L.nb_last_error(None) == b"invalid handle"
