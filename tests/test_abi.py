"""The C-ABI library loads and exports every symbol the headers declare (CPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in ("z2dqmc.h", "z2dqmc_test.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:int|const char \*)\s*\*?\s*(z2_\w+)\s*\(", txt, re.M)
    return names


def test_headers_declare_entry_points():
    names = _declared()
    assert "z2_ctx_create" in names and "z2_sweep" in names and len(names) > 20


def test_library_exports_declared_symbols():
    from z2_gauge_zfy_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libz2dqmc.so not built")
    L = _lib.lib()
    for n in _declared():
        assert hasattr(L, n), n
    assert set(_declared()) == set(_lib.exported_symbols())


def test_invalid_arguments_are_errors_not_aborts():
    from z2_gauge_zfy_amd import _lib
    import z2_gauge_zfy_amd as z
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libz2dqmc.so not built")
    with pytest.raises(z.Z2Error):  # odd L is rejected before any GPU call
        z.Engine(3, 3, 10, 5, [z.ChainSpec()])
    with pytest.raises(z.Z2Error):  # Ltrot not a multiple of Nwrap
        z.Engine(4, 4, 10, 3, [z.ChainSpec()])
