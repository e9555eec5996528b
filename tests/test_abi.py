"""CPU checks of the C-ABI library: it loads without a GPU and exports every
symbol include/rbchem.h declares (no compute calls here)."""
import os
import re

import pytest

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rbchem.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*\*?\s*(rb_\w+)\s*\(", text, re.M)))


def test_header_declares_the_documented_entry_points():
    from reprobus_amd import _lib

    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from reprobus_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("librbchem.so is not built (run __graft_entry__.build())")
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_struct_layouts_match_header():
    import ctypes as C

    from reprobus_amd import _lib

    assert C.sizeof(_lib.Grid) == 5 * 4
    assert C.sizeof(_lib.Config) == 4 * 4
    assert C.sizeof(_lib.Stats) == 8 * 7
