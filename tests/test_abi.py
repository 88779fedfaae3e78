"""The C-ABI library loads and exports every symbol that include/ddgrs.h declares (no GPU needed:
no compute call is made)."""
import ctypes
import os
import re

from conftest import ROOT, load_package

HEADER = os.path.join(ROOT, "include", "ddgrs.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"DDGRS_API\s+\w+\s+(ddgrs_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "ddgrs_ctx_create" in syms and "ddgrs_exact_energy_and_differential" in syms


def test_library_exports_every_declared_symbol():
    pkg = load_package()
    lib = ctypes.CDLL(pkg.library_path())
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    pkg = load_package()
    missing = [s for s in declared_symbols() if s not in pkg.SIGNATURES]
    assert not missing, missing
