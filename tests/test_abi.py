"""The C-ABI library loads and exports every entry point include/opem.h declares (no GPU needed)."""
import os
import re

from openparem_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "opem.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(opem_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SYMBOLS)


def test_create_without_gpu_fails_cleanly():
    """opem_create reports a runtime error code (never crashes) when no device is usable here."""
    import ctypes as C

    L = _lib.lib()
    ctx = C.c_void_p()
    rc = L.opem_create(C.byref(ctx), 9999)
    assert rc != 0
    assert L.opem_calculate_s(0, None, None, None, None, 0, None) < 0
