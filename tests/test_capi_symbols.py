"""The C-ABI library loads and exports every entry point include/vcp3d_amd.h
declares (no GPU calls)."""
import os
import re

from vcp3d_open_amd import solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vcp3d_amd.h")).read()
    return sorted(set(re.findall(r"\b(vcp_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(solver.EXPORTS)


def test_library_exports_all_symbols():
    L = solver.load_library()
    for name in declared_symbols():
        assert hasattr(L, name), name
