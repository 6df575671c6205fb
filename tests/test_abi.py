"""The C ABI boundary: libqpp_amd.so loads and exports every function that
include/qpp_amd.h declares; the C++ drop-in header compiles against it (CPU),
and the drop-in check program passes against the oracle (GPU)."""
import os
import re
import subprocess

import pytest

import qpp_amd as qa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qpp_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s+(qs_\w+)\s*\(", text, re.M)))


def exported_symbols(path):
    out = subprocess.check_output(["nm", "-D", "--defined-only", path], text=True)
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_header_declares_the_entry_points():
    fns = declared_functions()
    for f in ("qs_apply", "qs_apply_ctrl", "qs_apply_ctrl_fan", "qs_apply_kraus", "qs_ptrace",
              "qs_ptranspose", "qs_syspermute", "qs_sample", "qs_measure_seq",
              "qs_engine_execute", "qs_program_run"):
        assert f in fns


def test_library_exports_every_declared_symbol():
    lib = qa.lib()  # loads libqpp_amd.so (no device call)
    exported = exported_symbols(lib._name)
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    for f in declared_functions():
        assert hasattr(lib, f)


def test_last_error_without_device():
    lib = qa.lib()
    assert isinstance(lib.qs_last_error(), bytes)


def _build_dropin(out):
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"),
           "-o", out, os.path.join(ROOT, "tests", "cpp", "dropin_check.cpp"),
           "-L", os.path.join(ROOT, "qpp_amd"), "-lqpp_amd",
           "-L", os.path.join(ROOT, "oracle"), "-loracle",
           "-Wl,-rpath," + os.path.join(ROOT, "qpp_amd"),
           "-Wl,-rpath," + os.path.join(ROOT, "oracle")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_dropin_header_compiles(tmp_path):
    _build_dropin(str(tmp_path / "dropin"))


@pytest.mark.gpu
def test_dropin_check_program(tmp_path):
    exe = str(tmp_path / "dropin")
    _build_dropin(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
