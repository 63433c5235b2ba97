"""The C++ host API (include/bvhtracer.hpp): it compiles on its own, and the
reference-style test program built from tests/cpp passes on the GPU."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(HERE, "cpp", "test_reference_api")


def test_header_compiles_standalone():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
                        "-x", "c++", os.path.join(ROOT, "include", "bvhtracer.hpp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cpp_reference_tests_pass_on_gpu():
    assert os.path.exists(EXE), "tests/cpp/test_reference_api is built by __graft_entry__.build()"
    r = subprocess.run([EXE, os.path.join(HERE, "golden")], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS ") >= 30 and "FAIL" not in r.stdout
