"""Builds and runs tests/cpp/test_host.cpp against liblyap.so (CPU only), and
checks the chaos_lyap driver's start-up and config error paths."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "chaos_oneapi_amd")
CLI = os.path.join(PKG, "chaos_lyap")


def test_host_unit(tmp_path):
    exe = tmp_path / "test_host"
    cmd = [shutil.which("hipcc") or "/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-x", "c++",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(PKG, "csrc", "host"), os.path.join(ROOT, "tests", "cpp", "test_host.cpp"),
           "-o", str(exe), "-L" + PKG, "-llyap", "-Wl,-rpath," + PKG]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")


def run_cli(cwd, *args):
    env = dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    return subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=120, env=env)


def test_cli_lists_models_and_reports_missing_config(tmp_path):
    r = run_cli(tmp_path)
    assert r.returncode == 1
    assert 'Model[0]: "henon"  Dimension: 2 Number of parameters: 2' in r.stdout
    assert 'Model[2]: "lorenz"  Dimension: 3 Number of parameters: 3' in r.stdout
    assert "error: config file: config.json not found" in r.stdout


def test_cli_config_errors(tmp_path):
    (tmp_path / "config.json").write_text(
        '{"hardware": {"num_GPUs": 1}, "tool": {"lyapunov": {"samples_num": 1, "using_qr": false,'
        ' "using_var": false}}, "systems": []}')
    r = run_cli(tmp_path, "shimizu")
    assert r.returncode == 1 and "error: model: shimizu config not found" in r.stdout
    r = run_cli(tmp_path, "no_such_model")
    assert r.returncode == 1 and "Number of model is incorrect" in r.stdout
