"""ctypes mirror of include/merian_amd.h and the loader for the HIP library.

The library is built in-tree (merian_amd/libmerian_amd.so) by ``merian_amd.build``;
loading fails loudly when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libmerian_amd.so"

MRT_OK = 0
MRT_ERR_INVALID, MRT_ERR_DEVICE, MRT_ERR_BUDGET, MRT_ERR_UNSUPPORTED = -1, -2, -3, -4
MRT_ERR_NOT_READY, MRT_ERR_CANCELLED, MRT_ERR_ALLOC = -5, -6, -7
STATUS_NAMES = {0: "ok", -1: "invalid", -2: "device", -3: "budget", -4: "unsupported",
                -5: "not_ready", -6: "cancelled", -7: "alloc"}

# GeometryDataFlags (reference scene-data.slangh:13-41)
GEOM_PRETRANSFORMED = 0x1
GEOM_FLIP_FACING = 0x8
GEOM_INDEX_UINT32 = 0x40
GEOM_FLAT_SHADING = 0x80
GEOM_HAS_TANGENTS = 0x100
GEOM_USE_ENV_MAP = 0x200
GEOM_TWO_SIDED = 0x400

MAT_DIFFUSE = 0
MAT_OPENPBR = 1
PBR_TRANSMISSION = 0x1
PBR_CLEARCOAT = 0x2
PBR_SHEEN = 0x4
PBR_VOLUME = 0x8

ENV_NONE = 0
ENV_LATLONG = 1
ENV_EQUAL_AREA_OCT = 2
WRAP_REPEAT = 0
WRAP_CLAMP = 1
JITTER_NONE = 0
JITTER_R2 = 1
MODE_BSDF = 0
MODE_NEE = 1
MODE_GUIDED = 2
MODE_MCPG = 3

# numpy layouts of the POD arrays handed across the ABI
VERTEX_DTYPE = np.dtype([("position", "<f4", (3,)), ("encoded_normal", "<u4"),
                         ("uv", "<u2", (2,)), ("encoded_tangent", "<u4")])
assert VERTEX_DTYPE.itemsize == 24
GEOMETRY_DTYPE = np.dtype([("vertex_offset", "<u4"), ("index_offset", "<u4"),
                           ("primitive_count", "<u4"), ("material_id", "<u2"),
                           ("flags", "<u2"), ("instance_mask", "<u4")])
assert GEOMETRY_DTYPE.itemsize == 20
MATERIAL_DTYPE = np.dtype([("model", "<u4"), ("features", "<u4"), ("base_color", "<f4", (3,)),
                           ("opacity", "<f4"), ("metalness", "<f4"),
                           ("specular_alpha", "<f4", (2,)), ("specular_weight", "<f4"),
                           ("specular_ior", "<f4"), ("emission", "<f4", (3,)),
                           ("transmission_weight", "<f4"), ("transmission_color", "<f4", (3,)),
                           ("coat_weight", "<f4"), ("coat_alpha", "<f4"), ("coat_ior", "<f4"),
                           ("sheen_weight", "<f4"), ("sheen_color", "<f4", (3,)),
                           ("sheen_alpha", "<f4"), ("absorption", "<f4", (3,)),
                           ("_pad", "<f4", (3,))])
assert MATERIAL_DTYPE.itemsize == 128
RAY_DTYPE = np.dtype([("origin", "<f4", (3,)), ("tmin", "<f4"),
                      ("direction", "<f4", (3,)), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("triangle", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
PACKED_HIT_DTYPE = np.dtype([("instance_id", "<u4"), ("primitive_id", "<u4"),
                             ("geometry_index", "<u2"), ("instance_index", "<u2"),
                             ("barycentrics", "<u2", (2,))])
assert PACKED_HIT_DTYPE.itemsize == 16


class Env(C.Structure):
    _fields_ = [("type", C.c_int32), ("wrap", C.c_int32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("rgba", C.c_void_p), ("intensity", C.c_float),
                ("to_local", C.c_float * 9)]


class SceneDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_uint32),
                ("indices", C.c_void_p), ("index_count", C.c_uint32),
                ("geometries", C.c_void_p), ("geometry_count", C.c_uint32),
                ("materials", C.c_void_p), ("material_count", C.c_uint32),
                ("env", Env)]


class PbrtDesc(C.Structure):
    """mrt_pbrt_desc: a loaded pbrt scene as a host description plus its camera."""
    _fields_ = [("scene", SceneDesc), ("has_camera", C.c_int32), ("camera_from_file", C.c_int32),
                ("eye", C.c_float * 3), ("forward", C.c_float * 3), ("up", C.c_float * 3), ("vfov", C.c_float),
                ("aspect", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32),
                ("warning_count", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("U", C.c_float * 3), ("V", C.c_float * 3),
                ("W", C.c_float * 3), ("near_plane", C.c_float), ("far_plane", C.c_float),
                ("jitter", C.c_float * 2)]

    def as_dict(self):
        return {k: (list(getattr(self, k)) if k not in ("near_plane", "far_plane") else getattr(self, k))
                for k, _ in self._fields_}


class PtParams(C.Structure):
    _fields_ = [("spp", C.c_int32), ("max_path_length", C.c_int32),
                ("emission_on_primary", C.c_int32), ("demodulate_albedo", C.c_int32),
                ("instance_mask", C.c_uint32), ("mode", C.c_int32),
                ("record_signature", C.c_int32), ("count_traversal", C.c_int32)]


class PtStats(C.Structure):
    _fields_ = [("primary_segments", C.c_uint64), ("continuation_segments", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("ms_gbuffer", C.c_float), ("ms_trace", C.c_float),
                ("ms_shade", C.c_float), ("ms_total", C.c_float),
                ("trace_launches", C.c_uint32), ("_pad", C.c_uint32), ("node_visits", C.c_uint64),
                ("triangle_tests", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("max_paths", C.c_uint64), ("refill_lanes", C.c_uint32), ("grid_percent", C.c_uint32),
                ("bvh_max_leaf", C.c_uint32), ("bvh_traversal_cost", C.c_float), ("bvh_width", C.c_uint32),
                ("ms_per_frame", C.c_float), ("leaf_lanes", C.c_uint32), ("bvh8_order", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneRequest(C.Structure):
    _fields_ = [("camera", Camera), ("width", C.c_uint32), ("height", C.c_uint32), ("params", PtParams),
                ("time_budget_ms", C.c_float), ("_pad", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_slots", C.c_uint32), ("max_paths_per_job", C.c_uint32),
                ("use_graphs", C.c_int32), ("_pad", C.c_uint32)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("params", PtParams), ("width", C.c_uint32),
                ("height", C.c_uint32), ("frame", C.c_uint32), ("_pad", C.c_uint32),
                ("irradiance", C.c_void_p)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_CANCELLED, JOB_FAILED = 0, 1, 2, 3, 4
JOB_STATE_NAMES = {0: "queued", 1: "running", 2: "done", 3: "cancelled", 4: "failed"}


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("error", C.c_int32), ("segments", C.c_uint64),
                ("device_bytes", C.c_uint64), ("queued_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("cancelled", C.c_uint64),
                ("failed", C.c_uint64), ("segments", C.c_uint64), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("max_ms", C.c_double), ("mean_ms", C.c_double),
                ("wall_ms", C.c_double), ("peak_bytes", C.c_uint64), ("peak_in_flight", C.c_uint32),
                ("graph_instantiations", C.c_uint32), ("graph_updates", C.c_uint32), ("_pad", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("_")}


class GuideConfig(C.Structure):
    _fields_ = [("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3), ("seed", C.c_uint32),
                ("log2_table", C.c_uint32), ("base_resolution", C.c_uint32), ("max_resolution", C.c_uint32),
                ("train_batch", C.c_uint32), ("train_steps", C.c_uint32), ("max_records", C.c_uint32),
                ("learning_rate", C.c_float), ("guide_fraction", C.c_float), ("guided_bounces", C.c_uint32)]


GUIDE_LEVELS, GUIDE_FEATURES, GUIDE_WIDTH, GUIDE_LAYERS, GUIDE_OUTPUTS, GUIDE_LOBES = 8, 4, 64, 5, 40, 8


class McpgGridConfig(C.Structure):
    _fields_ = [("buffer_size", C.c_uint32), ("probe_count", C.c_uint32), ("locality_bits", C.c_uint32),
                ("stochastic_interpolation", C.c_uint32), ("tan_alpha_half", C.c_float), ("level_bias", C.c_float),
                ("distribution_dimension", C.c_float), ("_reserved", C.c_uint32)]


class McpgConfig(C.Structure):
    _fields_ = [("mc", McpgGridConfig), ("lc", McpgGridConfig), ("mc_samples", C.c_int32), ("p_guiding", C.c_float),
                ("sampling_type", C.c_int32), ("guiding_alpha_threshold", C.c_float),
                ("dir_guide_prior", C.c_float), ("missing_light_heuristic", C.c_int32),
                ("use_light_cache_tail", C.c_int32), ("lc_min_pdf", C.c_float), ("bsdf_support_only", C.c_int32)]


MCPG_MIS, MCPG_MIS_ROUGHNESS, MCPG_MAX_SAMPLES = 0, 1, 8
MCPG_STATE_DTYPE = np.dtype([("w_tgt", "<f4", 3), ("sum_w", "<f4"), ("w_cos", "<f4"), ("mv", "<u2", 3),
                             ("N", "<u2"), ("T", "<f4")])
LC_ENTRY_DTYPE = np.dtype([("irr", "<u2", 3), ("N", "<u2")])

assert C.sizeof(SceneDesc) == 128 and C.sizeof(Camera) == 64 and C.sizeof(PtParams) == 32
assert C.sizeof(ServerConfig) == 24 and C.sizeof(JobDesc) == 128 and C.sizeof(JobStatus) == 40
assert C.sizeof(ServerStats) == 104 and C.sizeof(GuideConfig) == 64
assert C.sizeof(McpgGridConfig) == 32 and C.sizeof(McpgConfig) == 100 and MCPG_STATE_DTYPE.itemsize == 32


def ptr(a: np.ndarray) -> C.c_void_p:
    """Address of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


_LIB = None

# (name, restype, argtypes) of every exported entry point of include/merian_amd.h
EXPORTS = [
    ("mrt_version", C.c_char_p, []),
    ("mrt_create", C.c_int32, [C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("mrt_destroy", C.c_int32, [C.c_void_p]),
    ("mrt_last_error", C.c_char_p, [C.c_void_p]),
    ("mrt_memory_in_use", C.c_uint64, [C.c_void_p]),
    ("mrt_synchronize", C.c_int32, [C.c_void_p]),
    ("mrt_scene_create", C.c_int32, [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    ("mrt_scene_destroy", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("mrt_scene_info", C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]),
    ("mrt_scene_create_triangles", C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                               C.c_uint32, C.c_int32, C.POINTER(C.c_void_p)]),
    ("mrt_camera_look_at", C.c_int32, [C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float), C.c_float, C.c_float, C.c_float,
                                       C.c_float, C.c_int32, C.c_uint32, C.POINTER(Camera)]),
    ("mrt_trace_closest", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    ("mrt_trace_closest_dev", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                          C.c_void_p, C.c_void_p]),
    ("mrt_trace_closest_noalloc_dev", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                                  C.c_void_p, C.c_void_p]),
    ("mrt_trace_count_dev", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                        C.POINTER(C.c_uint64)]),
    ("mrt_camera_rays_dev", C.c_int32, [C.c_void_p, C.POINTER(Camera), C.c_uint32, C.c_uint32,
                                        C.c_void_p, C.c_void_p]),
    ("mrt_gbuffer_dev", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.c_uint32,
                                    C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("mrt_gbuffer_albedo_dev", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mrt_pt_render_dev", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(PtParams),
                                      C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(PtStats),
                                      C.c_void_p]),
    ("mrt_render_frame", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(PtParams),
                                     C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(PtStats)]),
    ("mrt_server_create", C.c_int32, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    ("mrt_server_destroy", C.c_int32, [C.c_void_p]),
    ("mrt_server_last_error", C.c_char_p, [C.c_void_p]),
    ("mrt_job_bytes", C.c_int32, [C.c_void_p, C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
    ("mrt_job_submit", C.c_int32, [C.c_void_p, C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
    ("mrt_job_poll", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(JobStatus)]),
    ("mrt_job_wait", C.c_int32, [C.c_void_p, C.c_uint64, C.c_int32, C.POINTER(JobStatus)]),
    ("mrt_job_cancel", C.c_int32, [C.c_void_p, C.c_uint64]),
    ("mrt_server_stats_get", C.c_int32, [C.c_void_p, C.POINTER(ServerStats)]),
    ("mrt_server_stats_reset", C.c_int32, [C.c_void_p]),
    ("mrt_guide_layout", C.c_int32, [C.POINTER(GuideConfig), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("mrt_guide_create", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(GuideConfig), C.POINTER(C.c_void_p)]),
    ("mrt_guide_destroy", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("mrt_guide_normalisation", C.c_int32, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("mrt_guide_get_params", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mrt_guide_set_params", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mrt_guide_forward", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("mrt_guide_gradient", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_float)]),
    ("mrt_guide_train", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                    C.POINTER(C.c_float)]),
    ("mrt_guided_render_dev", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(PtParams),
                                          C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                          C.c_void_p, C.c_int32, C.POINTER(PtStats), C.c_void_p]),
    ("mrt_mcpg_default_config", None, [C.POINTER(McpgConfig)]),
    ("mrt_mcpg_create", C.c_int32, [C.c_void_p, C.POINTER(McpgConfig), C.POINTER(C.c_void_p)]),
    ("mrt_mcpg_destroy", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("mrt_mcpg_reset", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("mrt_mcpg_render_dev", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(PtParams),
                                        C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                        C.c_void_p, C.POINTER(PtStats), C.c_void_p]),
    ("mrt_mcpg_read_grids", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mrt_graph_apply_cli", C.c_int32, [C.c_char_p, C.c_int32, C.POINTER(C.c_char_p), C.c_char_p,
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("mrt_graph_resolve_includes", C.c_int32, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p),
                                               C.POINTER(C.c_void_p)]),
    ("mrt_graph_merge", C.c_int32, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("mrt_graph_store", C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("mrt_graph_cli_help", C.c_int32, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("mrt_string_free", None, [C.c_void_p]),
    ("mrt_pbrt_inspect", C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("mrt_pbrt_load_desc", C.c_int32, [C.c_char_p, C.POINTER(C.POINTER(PbrtDesc)), C.POINTER(C.c_void_p)]),
    ("mrt_pbrt_desc_free", None, [C.POINTER(PbrtDesc)]),
    ("mrt_pbrt_scene_create", C.c_int32, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("mrt_graph_run_main", C.c_int32, [C.c_int32, C.POINTER(C.c_char_p)]),
    ("mrt_alloc_dev", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("mrt_free_dev", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("mrt_download", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("mrt_upload", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("mrt_accumulate_dev", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    ("merian_node_plugin_abi_version", C.c_uint32, []),
    ("merian_register_nodes", None, [C.c_void_p]),
    ("mrt_scene_get_tuning", C.c_int32, [C.c_void_p, C.POINTER(Tuning)]),
    ("mrt_scene_set_tuning", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Tuning)]),
    ("mrt_tune", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(TuneRequest), C.POINTER(Tuning)]),
]


def load_library(path: os.PathLike | None = None) -> C.CDLL:
    """Load libmerian_amd.so; raises if it has not been built (no fallback path)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None and (PKG_DIR / "csrc").is_dir():
        # the library embeds a hash of the sources it was built from: a stale build fails loudly
        from .build import source_hash
        version = lib.mrt_version().decode()
        if not version.endswith("sources " + source_hash()):
            raise RuntimeError(f"{p} ({version}) was built from other sources: run __graft_entry__.build()")
    if path is None:
        _LIB = lib
    return lib
