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


def test_tune_bluestein_arguments():
    """csipm_tune_bluestein accepts only convolution lengths that serve the
    line length (2L - 1 <= M <= 4L, factorable) and 0 (remove); host-only."""
    from csipm import _lib
    from csipm import tune
    L = _lib.lib()
    assert L.csipm_tune_bluestein(111, 240) == 0       # 16 * 15
    assert L.csipm_tune_bluestein(111, 0) == 0
    assert L.csipm_tune_bluestein(111, 220) != 0       # below 2L - 1
    assert L.csipm_tune_bluestein(111, 222) != 0       # 2 * 3 * 37 does not factor
    assert L.csipm_tune_bluestein(111, 448) != 0       # above 4L
    assert L.csipm_tune_bluestein(2, 4) != 0
    tune.load(L)                                        # the committed table again
    # the tuner's factorisation agrees with the library's on every candidate
    for m in range(221, 445):
        ok = L.csipm_tune_bluestein(111, m) == 0
        assert ok == (tune.radix_stages(m) is not None), m
    L.csipm_tune_bluestein(111, 0)
    tune.load(L)
