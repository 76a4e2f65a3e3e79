"""The C++ operator surface (include/dgtd_hip.hpp) on the GPU: a program
linked against libdgtd_hip.so and the oracle checks Mult and an RK4
Solver::run against the oracle, the assembled variant (with and without
conductivity), and the three variants on a triangle mesh against each other
(tests/cpp/test_cpp_surface.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(__file__), "cpp", "_bin", "test_cpp_surface")


def test_cpp_surface_matches_oracle():
    assert os.path.exists(BIN), "build it first: make -C tests/cpp (part of __graft_entry__.build)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "OK" in r.stdout
