"""The C-ABI library loads and exports every symbol include/vortexhip.h declares (no GPU needed)."""
import ctypes as C
import os
import re

from vphip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "vortexhip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(vp_\w+)\(", src, re.M)))


def test_header_matches_binding_list():
    assert declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    lib = C.CDLL(_lib.LIBPATH)
    for name in declared():
        assert hasattr(lib, name), name
