"""Serving-loop admission policy and latency percentiles on the CPU (tests/native/test_serve_policy.cpp)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent


def test_serve_policy_native(tmp_path):
    exe = tmp_path / "test_serve_policy"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", str(ROOT / "native" / "test_serve_policy.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
