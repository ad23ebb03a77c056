"""CPU-side checks of the boundary: the library loads and exports exactly the
entry points include/k2g.h declares (no GPU call is made)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "k2g.h")).read()
    return sorted(set(re.findall(r"\b(k2g_[a-z_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    import kraken2_amd as k2
    lib = ctypes.CDLL(k2.LIB_PATH)
    names = declared_functions()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(lib, n), f"libk2g.so does not export {n}"
    assert sorted(k2.EXPORTED_SYMBOLS) == names
