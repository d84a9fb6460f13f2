"""bench.py's command line on the CPU: more than one GPU is refused before any GPU work, and
every leg can be switched off."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, env=env, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_distributed_launch_exits_with_a_message():
    env = dict(os.environ, WORLD_SIZE="2", RANK="0", LOCAL_RANK="0")
    r = _run("--gpus", "1", env=env)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_help_lists_every_leg_switch():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--no-cpu-baseline", "--no-guided", "--no-serving", "--no-ray-query",
                 "--no-small", "--retune", "--retune-rays"):
        assert flag in r.stdout, flag
