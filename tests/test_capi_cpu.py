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


def test_create_argument_errors_carry_messages_without_gpu():
    """Argument checks of baemm_create run before any HIP call; the message of
    a failed creation is readable through baemm_last_error(NULL)."""
    import numpy as np

    from baemm_amd.meshes import icosphere

    L = _lib.load()
    V, T = icosphere(1)
    bad = np.ascontiguousarray(T + V.shape[0])
    h = ctypes.c_void_p()
    st = L.baemm_create(V.ctypes.data, V.shape[0], bad.ctypes.data, bad.shape[0], None, 0, 0, ctypes.byref(h))
    assert st == _lib.ERR_ARG and not h.value
    assert b"triangle index" in L.baemm_last_error(None)
    st = L.baemm_create(V.ctypes.data, V.shape[0], T.ctypes.data, T.shape[0], None, -1, 0, ctypes.byref(h))
    assert st == _lib.ERR_ARG
    assert b"measurement" in L.baemm_last_error(None)


def test_header_constants_match_binding():
    """The status codes, element types, precision modes and wave types the
    binding uses are the values include/baemm.h defines."""
    import re
    text = (ROOT / "include" / "baemm.h").read_text()
    consts = {m.group(1): int(m.group(2)) for m in re.finditer(r"\b(BAEMM_[A-Z0-9_]+)\s*=\s*(-?\d+)", text)}
    assert consts["BAEMM_PRECISION_FP32"] == _lib.PRECISION_FP32
    assert consts["BAEMM_PRECISION_MIXED"] == _lib.PRECISION_MIXED
    assert consts["BAEMM_PRECISION_FP32_MFMA"] == _lib.PRECISION_FP32_MFMA
    assert consts["BAEMM_C128"] == _lib.C128 and consts["BAEMM_C64"] == _lib.C64
    assert consts["BAEMM_OK"] == _lib.OK and consts["BAEMM_ERR_ARG"] == _lib.ERR_ARG
    assert consts["BAEMM_ERR_NOCONV"] == _lib.ERR_NOCONV
    assert consts["BAEMM_SERVE_FIFO"] == _lib.SERVE_FIFO
    assert consts["BAEMM_SERVE_SHORTEST_FIRST"] == _lib.SERVE_SHORTEST_FIRST
