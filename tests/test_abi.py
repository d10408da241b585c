"""The C-ABI boundary: the product library loads without a GPU, exports every
symbol that include/*.h declares, has the 240-byte tEmDee layout, and the
reference's own host programs compile unchanged against include/ and link
to it (no GPU calls are made here)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "emdee_amd", "libemdee.so")
REF = "/root/reference"


def declared_symbols():
    names = []
    for h in ("emdee.h", "emdee_gpu.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"\b(EmDee(?:GPU)?_\w+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first: __graft_entry__.build()"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_symbols()) >= 54


def test_oracle_exports_the_reference_abi(oracle):
    names = [n for n in declared_symbols() if n.startswith("EmDee_")]
    missing = [n for n in names if not hasattr(oracle.lib, n)]
    assert not missing, missing


def test_struct_layout():
    from emdee_amd.emdee import tEmDee
    assert ctypes.sizeof(tEmDee) == 240
    off = {f: getattr(tEmDee, f).offset for f in ("Builds", "Time", "Energy", "Kinetic", "Virial",
                                                   "BodyVirial", "DoF", "RotDoF", "Data", "Options")}
    assert off == {"Builds": 0, "Time": 8, "Energy": 40, "Kinetic": 104, "Virial": 192, "BodyVirial": 200,
                   "DoF": 208, "RotDoF": 212, "Data": 216, "Options": 224}


def test_c_header_layout_matches(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text('#include "emdee.h"\n#include <stddef.h>\n#include <stdio.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(tEmDee), offsetof(tEmDee,Kinetic),'
                   'offsetof(tEmDee,Virial), offsetof(tEmDee,Data), offsetof(tEmDee,Options));return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["240", "104", "192", "216", "224"]


@pytest.mark.skipif(not os.path.exists(REF + "/src/testc.c"), reason="reference sources not present")
def test_reference_c_host_compiles_unchanged(tmp_path):
    exe = tmp_path / "testc"
    r = subprocess.run(["gcc", "-std=gnu99", "-I", os.path.join(ROOT, "include"), REF + "/src/testc.c",
                        "-o", str(exe), "-L", os.path.dirname(LIB), "-lemdee", "-lm",
                        "-Wl,-rpath," + os.path.dirname(LIB)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not (os.path.exists(REF + "/test/test_pair_lj_cut.f90")
                         and (shutil.which("amdflang") or os.path.exists("/opt/rocm/bin/amdflang"))),
                    reason="reference tests or amdflang absent")
def test_reference_fortran_host_compiles_unchanged(tmp_path):
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    common = REF + "/test/common"
    r = subprocess.run([fc, "-cpp", "-c", common + "/mConfig.f90", "-J", str(tmp_path), "-o", str(tmp_path / "mConfig.o")],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    exe = tmp_path / "test_pair_lj_cut"
    r = subprocess.run([fc, "-cpp", "-I", os.path.join(ROOT, "include"), "-I", REF + "/test", "-J", str(tmp_path),
                        REF + "/test/test_pair_lj_cut.f90", str(tmp_path / "mConfig.o"), "-o", str(exe),
                        "-L", os.path.dirname(LIB), "-lemdee", "-Wl,-rpath," + os.path.dirname(LIB)],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
