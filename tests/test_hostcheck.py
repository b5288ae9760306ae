"""Host-side library code on the CPU, built with g++ against the library's own
headers and run: re-validation (path_ok, metal-route_amd/csrc/hostcheck.h;
tests/hostcheck_cases.cpp: simple-path detection over long paths,
search-cost equality, adjacency and via rules, windows, obstacles and pads,
board bounds), the shared host workers (hostpool.h;
tests/hostpool_cases.cpp) and the persistent form's work-queue protocol
(workqueue.h; tests/workqueue_cases.cpp, a CPU model with worker threads).
Each also runs under the host sanitizers: AddressSanitizer with
UndefinedBehaviorSanitizer, and ThreadSanitizer for the threaded code."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


NEEDS = pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                           reason="needs g++ and the ROCm headers")


def build_and_run(tmp_path, name, flags=()):
    exe = tmp_path / (name + "".join(f.replace("=", "_").replace("-", "") for f in flags))
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-pthread", *flags, "-I/opt/rocm/include",
                    "-D__HIP_PLATFORM_AMD__", os.path.join(HERE, name + ".cpp"), "-o", str(exe)],
                   check=True, timeout=180)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


@NEEDS
def test_hostcheck_cases(tmp_path):
    build_and_run(tmp_path, "hostcheck_cases")


@NEEDS
def test_hostpool_cases(tmp_path):
    build_and_run(tmp_path, "hostpool_cases")


@NEEDS
def test_workqueue_cases(tmp_path):
    build_and_run(tmp_path, "workqueue_cases")


SAN_ASAN = ("-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer")
SAN_TSAN = ("-fsanitize=thread",)


@NEEDS
@pytest.mark.parametrize("name", ["hostcheck_cases", "hostpool_cases", "workqueue_cases"])
def test_cases_under_asan_ubsan(tmp_path, name):
    build_and_run(tmp_path, name, SAN_ASAN)


@NEEDS
@pytest.mark.parametrize("name", ["hostpool_cases", "workqueue_cases"])
def test_threaded_cases_under_tsan(tmp_path, name):
    build_and_run(tmp_path, name, SAN_TSAN)
