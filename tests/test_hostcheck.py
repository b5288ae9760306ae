"""Host-side library code on the CPU, built with g++ against the library's own
headers and run: re-validation (path_ok, metal-route_amd/csrc/hostcheck.h;
tests/hostcheck_cases.cpp: simple-path detection over long paths,
search-cost equality, adjacency and via rules, windows, obstacles and pads,
board bounds), the shared host workers (hostpool.h;
tests/hostpool_cases.cpp) and the persistent form's work-queue protocol
(workqueue.h; tests/workqueue_cases.cpp, a CPU model with worker threads);
the oracle's C code (oracle/mrcpu.c; tests/oracle_cases.c, its searches,
fields and threaded batches against one another); and the SimpleRouteJson
producer (metal-route_amd/csrc/srj.cpp; tests/srj_cases.cpp over the committed
fixtures, their prefixes and corrupted copies).
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


def build_and_run_c(tmp_path, flags):
    exe = tmp_path / ("oracle_cases" + "".join(f.replace("=", "_").replace("-", "").replace(",", "") for f in flags))
    subprocess.run(["gcc", "-O1", "-g", "-std=c11", *flags, os.path.join(HERE, "oracle_cases.c"),
                    os.path.join(HERE, "..", "oracle", "mrcpu.c"), "-lm", "-lpthread", "-o", str(exe)],
                   check=True, timeout=180)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr


NEEDS_GCC = pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")


@NEEDS_GCC
@pytest.mark.parametrize("flags", [(), SAN_ASAN, SAN_TSAN], ids=["plain", "asan_ubsan", "tsan"])
def test_oracle_cases(tmp_path, flags):
    build_and_run_c(tmp_path, flags)


@NEEDS
@pytest.mark.parametrize("flags", [(), SAN_ASAN], ids=["plain", "asan_ubsan"])
def test_srj_cases(tmp_path, flags):
    import glob

    exe = tmp_path / ("srj_cases" + "".join(f.replace("=", "_").replace("-", "").replace(",", "") for f in flags))
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", *flags, os.path.join(HERE, "srj_cases.cpp"),
                    os.path.join(HERE, "..", "metal-route_amd", "csrc", "srj.cpp"), "-o", str(exe)],
                   check=True, timeout=300)
    files = sorted(glob.glob(os.path.join(HERE, "golden", "srj", "*.json")))
    r = subprocess.run([str(exe), *files], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
