"""CPU-side checks of the C ABI: the built library loads and exports every
function declared in include/baemm.h; the Python binding covers them all."""
import ctypes
import re

import pytest

from baemm_amd import _lib
from conftest import ROOT


def _declared():
    text = (ROOT / "include" / "baemm.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(baemm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = _declared()
    assert "baemm_create" in names and "baemm_apply_boundary" in names
    assert len(names) > 30


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libbaemm_hip.so not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_declared()) == set(_lib.SIGNATURES)


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libbaemm_hip.so not built")
def test_null_handle_is_rejected_without_gpu():
    L = _lib.load()
    assert L.baemm_version() >= 10000
    assert L.baemm_destroy(None) == _lib.ERR_ARG
    assert L.baemm_vertex_count(None) == -1
    assert L.baemm_use_diagonal(None, 1) == _lib.ERR_ARG
    h = ctypes.c_void_p()
    assert L.baemm_create(None, 0, None, 0, None, 0, 0, ctypes.byref(h)) == _lib.ERR_ARG
