"""The C-ABI library loads (no GPU needed) and exports every symbol that
include/csipm.h declares, with the argument count the ctypes binding uses."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "csipm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(csipm_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding():
    from csipm import _lib
    assert declared() == sorted(_lib.SIGNATURES)


def test_library_exports_all_symbols():
    from csipm import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libcsipm.so is not built (run __graft_entry__.build())")
    L = _lib.lib()
    for name in declared():
        assert hasattr(L, name), name
    assert L.csipm_version().decode().count(".") == 2
