"""The host C++ (mesh set-up, reference element, connectivity, case-file and
mesh readers, Global-operator assembly, initial-field projection) built with
AddressSanitizer and UndefinedBehaviorSanitizer (make -C dgtd_amd/csrc asan)
runs the host-side tests in a child process with the sanitizer runtime
preloaded; any report aborts it."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES = ["tests/test_host_setup.py", "tests/test_case_files.py", "tests/test_global_assembly.py"]


def _gcc_file(tool, name):
    return subprocess.run([tool, f"-print-file-name={name}"], capture_output=True, text=True,
                          check=True).stdout.strip()


def test_host_code_under_asan_and_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "dgtd_amd", "csrc"), "asan"], check=True)
    lib = os.path.join(ROOT, "dgtd_amd", "lib_asan", "libdgtd_hip.so")
    asan, stdcxx = _gcc_file("gcc", "libasan.so"), _gcc_file("g++", "libstdc++.so.6")
    if not (os.path.exists(asan) and os.path.exists(stdcxx)):
        pytest.skip("no sanitizer runtime in this toolchain")
    log = str(tmp_path / "san")
    env = dict(os.environ, DGTD_LIB=lib, LD_PRELOAD=f"{asan} {stdcxx}",
               ASAN_OPTIONS=f"detect_leaks=0:abort_on_error=1:log_path={log}",
               UBSAN_OPTIONS=f"halt_on_error=1:print_stacktrace=1:log_path={log}")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + FILES,
                       cwd=ROOT, env=env, capture_output=True, text=True)
    reports = [f for f in os.listdir(tmp_path) if f.startswith("san")]
    detail = "".join(open(tmp_path / f).read()[:4000] for f in reports)
    assert r.returncode == 0 and not reports, (r.stdout[-2000:], detail)
