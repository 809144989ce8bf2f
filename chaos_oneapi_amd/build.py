"""Build the gfx950 library (liblyap.so) and the command-line driver in-tree.

    python -m chaos_oneapi_amd.build          # or build() from Python

Every HIP/C++ source under csrc/ is compiled by hipcc for gfx950 with
-ffp-contract=off (the FP64 path is bit-reproducible against the CPU oracle)
and with packed FP32 instructions disabled.  Objects go to _build/, the
library to chaos_oneapi_amd/liblyap.so, the driver to chaos_oneapi_amd/chaos_lyap,
and the vector-ALU peak probe (csrc/probe/valu_peak.hip, read by bench.py's
roofline) to chaos_oneapi_amd/valu_peak.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(CSRC, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "liblyap.so")
CLI = os.path.join(PKG, "chaos_lyap")
PROBE_SRC = os.path.join(CSRC, "probe", "valu_peak.hip")
PROBE = os.path.join(PKG, "valu_peak")

ARCH = os.environ.get("CHAOS_LYAP_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + HOST,
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


HOSTONLY = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def _sources():
    # csrc/*.hip and csrc/*.cpp see the device headers: compiled as HIP
    dev = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))
    host = []
    if os.path.isdir(HOST):
        host += sorted(os.path.join(HOST, f) for f in os.listdir(HOST)
                       if f.endswith(".cpp") and f != "main.cpp")
    return dev, host


def _compile(src: str, device: bool, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + (["-x", "hip"] + DEVICE if device else HOSTONLY)
    cmd += ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source and link liblyap.so and chaos_lyap; returns the library path."""
    os.makedirs(BUILD, exist_ok=True)
    dev, host = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, True, verbose) for s in dev]
        futs += [ex.submit(_compile, s, False, verbose) for s in host]
        objs = [f.result() for f in futs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main):
        cmd = [HIPCC] + COMMON + HOSTONLY + [main, "-o", CLI + ".tmp",
                                  "-L" + PKG, "-llyap", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver link failed\n{r.stdout}\n{r.stderr}")
        os.replace(CLI + ".tmp", CLI)
    if os.path.exists(PROBE_SRC):
        cmd = [HIPCC, "-O3", "-std=c++17", "-x", "hip"] + DEVICE + [PROBE_SRC, "-o", PROBE + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"probe build failed\n{r.stdout}\n{r.stderr}")
        os.replace(PROBE + ".tmp", PROBE)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
