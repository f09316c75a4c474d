"""The C-ABI libraries load on a machine without a GPU and export every
function declared in include/tdfstat_hip.h and include/tdfstat_search.h (no
compute calls here)."""
import ctypes
import os
import re

from tdfstat_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="tdfstat_hip.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tdf_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_api():
    names = _declared()
    assert "tdf_job_core" in names and "tdf_ctx_create" in names
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_strerror_without_gpu():
    lib = L.load()
    assert lib.tdf_strerror(L.TDF_OK) == b"ok"
    assert b"Too many triggers" in lib.tdf_strerror(-4)


def test_search_library_exports_every_symbol():
    names = _declared("tdfstat_search.h")
    assert set(names) == set(L.SEARCH_SIGNATURES), set(names) ^ set(L.SEARCH_SIGNATURES)
    lib = L.load_search()
    for name in names:
        assert hasattr(lib, name), name


def test_memory_ceilings_rejects_bad_sizes_without_gpu():
    lib = L.load()
    g = (ctypes.c_double * 6)()
    for nbytes, reps in ((0, 1), (17, 1), (1 << 20, 0)):
        assert lib.tdf_memory_ceilings(nbytes, reps, g) == L.TDF_ERR_ARG
    assert lib.tdf_memory_ceilings(1 << 20, 1, None) == L.TDF_ERR_ARG
