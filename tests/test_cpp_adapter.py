"""The VecGeom-shaped C++ adapter (include/vgm/vecgeom_compat.h): it compiles
against the C ABI on the CPU, and the test program written against it passes
on the GPU (closed-form box queries, a two-volume navigation, errors as
exceptions)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "test_compat")


def test_adapter_compiles():
    src = os.path.join(ROOT, "tests", "cpp", "test_compat.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_adapter_program_runs():
    assert os.path.exists(PROG), "build the program first (__graft_entry__.build or make -C tests/cpp)"
    print("[adapter] running tests/cpp/test_compat", flush=True)
    r = subprocess.run([PROG], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().startswith("ok")
