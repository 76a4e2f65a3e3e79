"""include/dgtd_hip.hpp and include/dgtd_hip.h compile as C++17 and C99."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_headers_compile():
    inc = os.path.join(ROOT, "include")
    src = os.path.join(ROOT, "tests", "cpp", "test_cpp_surface.cpp")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", inc, src], check=True)
    subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Werror", "-x", "c", os.path.join(inc, "dgtd_hip.h")],
                   check=True)
