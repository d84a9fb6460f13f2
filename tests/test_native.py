"""Native CPU unit tests: the serving loop's admission policy (tests/native/test_serve_policy.cpp), the
8-wide quantised BVH builder (tests/native/test_bvh8.cpp) and the graph host's node lifecycle and events
(tests/native/test_graph_lifecycle.cpp)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent


def test_serve_policy_native(tmp_path):
    exe = tmp_path / "test_serve_policy"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", str(ROOT / "native" / "test_serve_policy.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_bvh8_structure_native(tmp_path):
    exe = tmp_path / "test_bvh8"
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", str(ROOT / "native" / "test_bvh8.cpp"),
                    str(ROOT.parent / "merian_amd" / "csrc" / "bvh.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_graph_lifecycle_native(tmp_path):
    exe = tmp_path / "test_graph_lifecycle"
    host = ROOT.parent / "merian_amd" / "host"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", str(ROOT / "native" / "test_graph_lifecycle.cpp"),
                    str(host / "graph.cpp"), str(host / "graph_description.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
