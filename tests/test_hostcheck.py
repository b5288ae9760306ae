"""Host-side library code on the CPU, built with g++ against the library's own
headers and run: re-validation (path_ok, metal-route_amd/csrc/hostcheck.h;
tests/hostcheck_cases.cpp: simple-path detection over long paths,
search-cost equality, adjacency and via rules, windows, obstacles and pads,
board bounds) and the shared host workers (hostpool.h;
tests/hostpool_cases.cpp)."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


NEEDS = pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                           reason="needs g++ and the ROCm headers")


def build_and_run(tmp_path, name):
    exe = tmp_path / name
    subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                    os.path.join(HERE, name + ".cpp"), "-o", str(exe)], check=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@NEEDS
def test_hostcheck_cases(tmp_path):
    build_and_run(tmp_path, "hostcheck_cases")


@NEEDS
def test_hostpool_cases(tmp_path):
    build_and_run(tmp_path, "hostpool_cases")
