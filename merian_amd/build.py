"""Builds merian_amd/libmerian_amd.so with hipcc for gfx950 (in-tree, no JIT cache).

Usage: python -m merian_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmerian_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["csrc/kernels.hip", "csrc/api.cpp", "csrc/bvh.cpp", "csrc/serve.cpp", "csrc/guide.hip",
           "csrc/guide_api.cpp", "csrc/tune.cpp", "csrc/mcpg_api.cpp", "csrc/noalloc.cpp", "host/graph_description.cpp", "host/graph_abi.cpp", "host/pbrt.cpp",
           "host/graph.cpp", "host/nodes.cpp", "host/graph_run.cpp"]
RUNNER = PKG / "merian-amd-graph-run"

# -ffp-contract=off: expressions round as written (the CPU oracle uses the same rule).
# -packed-fp32-ops off: no v_pk_*_f32 generation.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-fvisibility=hidden", "-ffp-contract=off",
          "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result"]


def source_hash() -> str:
    """SHA-256 (first 16 hex digits) over the library's sources, headers and this build script:
    the library embeds it (mrt_version) so a stale build is detected at load time."""
    import hashlib
    h = hashlib.sha256()
    files = sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp"), *CSRC.glob("*.h"), *(PKG / "host").glob("*.cpp"),
                    *(PKG / "host").glob("*.h"), PKG.parent / "include" / "merian_amd.h", Path(__file__).resolve()])
    for f in files:
        h.update(f.name.encode())
        h.update(f.read_bytes())
    return h.hexdigest()[:16]


def _compile(src: Path, defines: list[str] = ()) -> Path:
    obj = OBJ / (src.parent.name + "_" + src.name + ".o")
    cmd = [HIPCC, *COMMON, *defines, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src.name}")
    return obj


def build(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = [PKG / s for s in SOURCES]
    defines = [f'-DMRT_SOURCE_HASH="{source_hash()}"']
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, 16))) as ex:
        objs = list(ex.map(lambda s: _compile(s, defines), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs), "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    # the runner executable links the library by relative rpath
    main_obj = _compile(PKG / "host" / "graph_run_main.cpp")
    tmp = RUNNER.with_suffix(".tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-o", str(tmp), str(main_obj), f"-L{PKG}", "-l:libmerian_amd.so",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("runner link failed")
    os.replace(tmp, RUNNER)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    print(build(ap.parse_args().jobs))
