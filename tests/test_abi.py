"""The C ABI without a GPU: the library loads, exports every entry point include/merian_amd.h
declares (and nothing the header does not), struct layouts agree between the header and the
ctypes mirror, and device entry points fail with a status instead of crashing.
"""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from merian_amd import abi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "merian_amd.h"


def declared():
    text = HEADER.read_text()
    return set(re.findall(r"MRT_API\s+[\w\s\*]+?\b(\w+)\s*\(", text))


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_exactly_the_header():
    d = declared()
    assert len(d) >= 18
    assert d == exported(), (d ^ exported())
    assert {name for name, _, _ in abi.EXPORTS} == d


def test_ctypes_layouts_match_the_header(tmp_path):
    src = tmp_path / "layout_probe.c"
    exe = tmp_path / "layout_probe"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "include/merian_amd.h"\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\","
                   "sizeof(mrt_vertex),sizeof(mrt_geometry),sizeof(mrt_material),sizeof(mrt_env),"
                   "sizeof(mrt_scene_desc),sizeof(mrt_camera),sizeof(mrt_ray),sizeof(mrt_hit),"
                   "sizeof(mrt_packed_hit),sizeof(mrt_pt_params),sizeof(mrt_pt_stats),"
                   "offsetof(mrt_pt_stats,node_visits),sizeof(mrt_guide_config),"
                   "offsetof(mrt_guide_config,guided_bounces),sizeof(mrt_mcpg_grid_config),"
                   "sizeof(mrt_mcpg_config),offsetof(mrt_mcpg_config,lc_min_pdf));"
                   "printf(\"%zu %zu\\n\",sizeof(mrt_pbrt_desc),offsetof(mrt_pbrt_desc,warning_count));return 0;}\n")
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT), "-o", str(exe), str(src)], check=True)
    sizes = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    ours = [abi.VERTEX_DTYPE.itemsize, abi.GEOMETRY_DTYPE.itemsize, abi.MATERIAL_DTYPE.itemsize,
            C.sizeof(abi.Env), C.sizeof(abi.SceneDesc), C.sizeof(abi.Camera), abi.RAY_DTYPE.itemsize,
            abi.HIT_DTYPE.itemsize, abi.PACKED_HIT_DTYPE.itemsize, C.sizeof(abi.PtParams),
            C.sizeof(abi.PtStats), abi.PtStats.node_visits.offset, C.sizeof(abi.GuideConfig),
            abi.GuideConfig.guided_bounces.offset, C.sizeof(abi.McpgGridConfig), C.sizeof(abi.McpgConfig),
            abi.McpgConfig.lc_min_pdf.offset, C.sizeof(abi.PbrtDesc), abi.PbrtDesc.warning_count.offset]
    assert sizes == ours


def test_plugin_abi_and_version():
    lib = abi.load_library()
    assert lib.merian_node_plugin_abi_version() == 1
    assert b"gfx950" in lib.mrt_version()


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = abi.load_library()
    h = C.c_void_p()
    st = lib.mrt_create(0, 0, C.byref(h))
    assert st in (-2, -1) and not h.value
    assert lib.mrt_create(0, 0, None) == -1
    # null handles are rejected, not dereferenced
    assert lib.mrt_destroy(None) == -1
    assert lib.mrt_synchronize(None) == -1
    assert lib.mrt_scene_info(None, None, None, None) == -1
    cam = abi.Camera()
    f3 = C.c_float * 3
    assert lib.mrt_camera_look_at(f3(0, 0, 0), f3(0, 0, -1), f3(0, 1, 0), 1.0, 1.0, 0.1, 10.0, 7, 0, C.byref(cam)) == -1


def test_camera_is_host_side_and_matches_oracle():
    # mrt_camera_look_at is plain host arithmetic (camera.cpp:349-362); it must equal the oracle's
    from merian_amd import scenes
    from oracle import oracle as O
    lib = abi.load_library()
    for spec in (scenes.plain_small().camera, scenes.ray_query_camera(512)):
        for frame in (0, 3, 17):
            cam = abi.Camera()
            f3 = C.c_float * 3
            assert lib.mrt_camera_look_at(f3(*spec.eye), f3(*spec.target), f3(*spec.up), spec.vfov, spec.aspect,
                                          spec.near, spec.far, spec.jitter, frame, C.byref(cam)) == 0
            assert bytes(cam) == bytes(O.camera_look_at(spec, frame))


def test_library_is_built_from_these_sources():
    # mrt_version carries the hash of the sources the library was built from (merian_amd/build.py);
    # load_library refuses a library built from other sources
    from merian_amd.build import source_hash
    lib = abi.load_library()
    assert lib.mrt_version().decode().endswith("sources " + source_hash())


def test_stream_argument_mapping():
    # None selects the context's own stream (NULL); HIP stream handles pass through, except the
    # null stream (handle 0, torch's default stream), which is passed as hipStreamLegacy
    from merian_amd.runtime import HIP_STREAM_LEGACY, _stream
    assert _stream(None) is None
    assert _stream(0).value == HIP_STREAM_LEGACY == 1
    assert _stream(0x7f00deadbeef).value == 0x7f00deadbeef
