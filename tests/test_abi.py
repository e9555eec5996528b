"""CPU checks of the C-ABI library: it loads without a GPU and exports every
symbol include/rbchem.h, include/rbmlp.h and include/rbio.h declare (no compute calls
here)."""
import os
import re

import pytest

from conftest import ROOT


def declared_symbols(header="rbchem.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|void|const char \*)\s*\*?\s*(rb_\w+)\s*\(",
                                 text, re.M)))


def test_header_declares_the_documented_entry_points():
    from reprobus_amd import _lib

    assert declared_symbols() == sorted(_lib.EXPORTS)
    assert declared_symbols("rbmlp.h") == sorted(_lib.MLP_EXPORTS)
    assert declared_symbols("rbio.h") == sorted(_lib.IO_EXPORTS)


def test_library_exports_every_declared_symbol():
    from reprobus_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("librbchem.so is not built (run __graft_entry__.build())")
    L = _lib.load()
    for name in declared_symbols() + declared_symbols("rbmlp.h") + declared_symbols("rbio.h"):
        assert hasattr(L, name), name
    assert L.rb_mlp_param_count() == 64 * 256 + 256 + 2 * (256 * 256 + 256) + 256 * 64 + 64


def test_struct_layouts_match_header():
    import ctypes as C

    from reprobus_amd import _lib

    assert C.sizeof(_lib.Grid) == 5 * 4
    assert C.sizeof(_lib.Config) == 9 * 4
    assert C.sizeof(_lib.Stats) == 8 * 7
    assert C.sizeof(_lib.DayConfig) == 4 * 4 + 3 * 8 + 2 * 4


def test_struct_sizes_match_the_compiled_header(tmp_path):
    """The ctypes mirrors have the sizes the C compiler gives the header's structs."""
    import ctypes as C
    import shutil
    import subprocess

    from reprobus_amd import _lib

    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "rbchem.h"\nint main(void) {\n'
                   '  printf("%zu %zu %zu %zu\\n", sizeof(rb_grid), sizeof(rb_config),\n'
                   '         sizeof(rb_stats_t), sizeof(rb_day_config));\n  return 0;\n}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got == [C.sizeof(_lib.Grid), C.sizeof(_lib.Config), C.sizeof(_lib.Stats),
                   C.sizeof(_lib.DayConfig)]
