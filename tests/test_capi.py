"""The C ABI library loads and exports every symbol include/tailor.h declares.

No compute call is made here (this runs without a GPU).
"""
import ctypes
import re

from conftest import ROOT


def declared_symbols():
    text = (ROOT / "include" / "tailor.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tailor_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "tailor_create" in syms and "tailor_assemble" in syms
    assert len(syms) >= 10


def test_library_exports_every_declared_symbol():
    import tailor_amd
    lib = tailor_amd.library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_library_is_gfx950_hip_code():
    # the shared object embeds a gfx950 code object
    data = (ROOT / "tailor_amd" / "libtailor.so").read_bytes()
    assert b"gfx950" in data
    assert isinstance(ctypes.CDLL(str(ROOT / "tailor_amd" / "libtailor.so")), ctypes.CDLL)
