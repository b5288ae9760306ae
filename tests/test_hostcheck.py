"""Host re-validation (path_ok, metal-route_amd/csrc/hostcheck.h) on the CPU:
tests/hostcheck_cases.cpp is built with g++ against the library's own header
and run.  Simple-path detection over long paths, search-cost equality,
adjacency and via rules, windows, obstacles and pads, board bounds."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the ROCm headers")
def test_hostcheck_cases(tmp_path):
    exe = tmp_path / "hostcheck_cases"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                    os.path.join(HERE, "hostcheck_cases.cpp"), "-o", str(exe)], check=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
