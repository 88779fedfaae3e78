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


def _c_layout(tmp_path, structs):
    """sizeof and field offsets of the header's structs, from gcc."""
    import subprocess

    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "ddgrs.h"', "int main(void) {"]
    for name, fields in structs.items():
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'  printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def test_struct_layouts_match_binding(tmp_path):
    pkg = load_package()
    pairs = {
        "ddgrs_error": pkg._Err,
        "ddgrs_saddle_options": pkg._SaddleOptionsC,
        "ddgrs_gmres_report": pkg._GmresReportC,
        "ddgrs_step_options": pkg._StepOptionsC,
        "ddgrs_step_report": pkg._StepReportC,
        "ddgrs_ccd_options": pkg._CcdOptionsC,
        "ddgrs_ccd_report": pkg._CcdReportC,
        "ddgrs_remesh_options": pkg._RemeshOptionsC,
        "ddgrs_remesh_report": pkg._RemeshReportC,
    }
    c = _c_layout(tmp_path, {k: [f for f, _ in v._fields_] for k, v in pairs.items()})
    for name, cls in pairs.items():
        assert c[name] == ctypes.sizeof(cls), name
        for f, _ in cls._fields_:
            assert c[f"{name}.{f}"] == getattr(cls, f).offset, (name, f)


def _rust_stub():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return text[text.index("```rust\n#![allow"):]


def test_rust_stub_structs_match_header():
    """The -sys crate stub in INTEGRATION.md declares every header struct with the same fields in
    the same order (the binding's ctypes structs, whose layout the test above checks against gcc)."""
    pkg = load_package()
    stub = _rust_stub()
    pairs = {
        "ddgrs_saddle_options": pkg._SaddleOptionsC,
        "ddgrs_gmres_report": pkg._GmresReportC,
        "ddgrs_step_options": pkg._StepOptionsC,
        "ddgrs_step_report": pkg._StepReportC,
        "ddgrs_ccd_options": pkg._CcdOptionsC,
        "ddgrs_ccd_report": pkg._CcdReportC,
        "ddgrs_remesh_options": pkg._RemeshOptionsC,
        "ddgrs_remesh_report": pkg._RemeshReportC,
    }
    for name, cls in pairs.items():
        body = re.search(r"pub struct " + name + r"\s*\{(.*?)\}", stub, re.S).group(1)
        fields = re.findall(r"pub (\w+):", body)
        assert fields == [f for f, _ in cls._fields_], name


def test_rust_stub_declares_every_entry_point():
    stub = _rust_stub()
    declared = set(re.findall(r"pub fn (ddgrs_\w+)\(", stub))
    missing = [s for s in declared_symbols() if s not in declared]
    assert not missing, missing
