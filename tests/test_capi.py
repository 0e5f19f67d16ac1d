"""The C-ABI library loads and exports every entry point include/sevp.h
declares (no GPU calls)."""
import ctypes
import os
import re

from musoc_solver_amd import sevp as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sevp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sevp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("sevp_plan_create", "sevp_solve", "sevp_repbir", "sevp_rep1", "sevp_pre",
                 "sevp_pre1", "sevp_matinv", "sevp_solve_refined_dev"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = S.lib()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert "gfx950" in S.version()


def test_fortran_shim_exports_reference_names():
    path = os.path.join(ROOT, "musoc_solver_amd", "lib", "libsevp_fortran.so")
    if not os.path.exists(path):
        import pytest
        pytest.skip("Fortran shim not built (no flang)")
    lib = ctypes.CDLL(path)
    for name in ("repbir_", "rep1_", "rep_", "repbir_period_", "pre_", "pre1_", "matinv_"):
        assert hasattr(lib, name), name
