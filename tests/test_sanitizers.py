"""AddressSanitizer + UBSan runs of the CPU code (no GPU): the oracle with its
self-test driver (every model and method), and the C++ host layer
(config/JSON parsing, the model registry, TaskLyapunov's host logic, the
output writer) with its unit tests.  Built with gcc/g++ for the host only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "chaos_oneapi_amd")
HOST = os.path.join(PKG, "csrc", "host")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="needs gcc/g++")


def test_oracle_under_asan_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "selftest-san",
                    f"SANDIR={tmp_path}"], check=True, capture_output=True)
    r = subprocess.run([str(tmp_path / "orc_selftest_san")], capture_output=True, text=True,
                       timeout=600, env=ENV)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "OK"


def test_host_layer_under_asan_ubsan(tmp_path):
    srcs = [os.path.join(HOST, f) for f in ("json.cpp", "tool_params.cpp", "model_factory.cpp",
                                            "tasklyapunov.cpp", "filehandler.cpp")]
    exe = tmp_path / "test_host_san"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-Wno-unused-result",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
           "-I" + HOST, *srcs, os.path.join(ROOT, "tests", "cpp", "test_host.cpp"), "-o", str(exe),
           "-L" + PKG, "-llyap", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath," + PKG,
           "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
