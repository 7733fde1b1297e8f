"""The C-ABI library loads and exports every symbol include/efg.h declares."""
import ctypes

from efanna_amd import _lib


def test_header_parses():
    syms = _lib.header_symbols()
    assert "efg_create" in syms and "efg_knn_search" in syms and len(syms) >= 30


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [s for s in _lib.header_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_wrapper_signatures_cover_header():
    assert set(_lib.header_symbols()) == set(_lib._SIGS), set(_lib.header_symbols()) ^ set(_lib._SIGS)


def test_error_path_without_gpu_work():
    L = _lib.lib()
    out = ctypes.c_void_p()
    rc = L.efg_create(None, 0, 0, 0, None, ctypes.byref(out))
    assert rc == -2
    assert b"invalid" in L.efg_last_error()
