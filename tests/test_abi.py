"""The C-ABI libraries load and export every entry point their headers declare
(no device calls: runs without a GPU)."""
import os
import re
import subprocess

import pytest

from gascoignelib_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("header,prefix,lib,sigs", [
    ("gascoigne_hip.h", "gh_", N.HIP_LIB, N.HIP_SIGNATURES),
    ("gh_structure.h", "ghs_", N.STRUCT_LIB, N.STRUCT_SIGNATURES),
])
def test_every_declared_symbol_is_exported_and_bound(header, prefix, lib, sigs):
    names = declared(header, prefix)
    assert len(names) > 10
    exp = exported(lib)
    missing = [n for n in names if n not in exp]
    assert not missing, f"{lib} does not export {missing}"
    unbound = [n for n in names if n not in sigs]
    assert not unbound, f"native.py has no signature for {unbound}"


def test_libraries_load():
    assert N.hip().gh_version().decode().startswith("gascoigne_hip")
    assert N.structure().ghs_last_error() is not None


def test_hip_library_is_gfx950():
    # the offload bundle of the library carries its target id
    data = open(N.HIP_LIB, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
