"""The C-ABI library loads and exports every entry point include/ld.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from lucky_droplets_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "ld.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ld_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_entry_points():
    names = declared()
    for n in ("ld_rfft", "ld_irfft", "ld_forcing", "ld_clip_zero", "ld_full_rhs", "ld_rk4", "ld_step",
              "ld_interp_cosine", "ld_exterp_cosine_scalar", "ld_prhs"):
        assert n in names


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libld.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == set(declared())


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libld.so not built")
def test_argument_errors_without_gpu():
    """Argument validation happens before any device call."""
    lib = _lib.load()
    assert lib.ld_version().startswith(b"ld ")
    assert lib.ld_rfft(None, None, None, 1) == _lib.LD_ERR_ARG
    assert "ld_ctx" in _lib.last_error()
    p = _lib.LdParams()
    p.N = 48
    p.precision = _lib.LD_F32
    ctx = ctypes.c_void_p()
    assert lib.ld_create(ctypes.byref(p), ctypes.byref(ctx)) == _lib.LD_ERR_ARG
    assert "N=48" in _lib.last_error()
    p.N = 64
    p.precision = 7
    assert lib.ld_create(ctypes.byref(p), ctypes.byref(ctx)) == _lib.LD_ERR_ARG
