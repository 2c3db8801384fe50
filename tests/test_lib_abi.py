"""The C-ABI library loads (no GPU needed) and exports every entry point the
public header declares."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "damask_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(dmk_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "dmk_mech_init" in names and "dmk_mech_solution" in names and len(names) >= 17


def test_library_exports_all_declared_symbols():
    from damask_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(_declared()) <= set(_lib.EXPORTS) | set(_declared())


def test_python_binding_lists_all_exports():
    from damask_amd import _lib
    assert set(_declared()) == set(_lib.EXPORTS)
