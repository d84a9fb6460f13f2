"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU oracle restates the reference's pt path (render_pt.slang and the Slang library it
imports). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker / CPU baseline; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from merian_amd import abi

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class BsdfCfg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("albedo", C.c_float * 3), ("alpha", C.c_float * 2), ("f0", C.c_float),
                ("eta", C.c_float), ("front_facing", C.c_int32), ("ior", C.c_float),
                ("mat", C.c_uint8 * 128)]


class GuideSpec(C.Structure):
    """Guide encoding layout: per-level grid resolutions, table entries, position normalisation."""
    _fields_ = [("res", C.c_uint32 * 8), ("table", C.c_uint32), ("offset", C.c_float * 3),
                ("scale", C.c_float * 3)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp, u32, u64, i32, f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_float
        sig = {
            "orc_pcg4d16": (u32, [u32, u32, u32, u32]),
            "orc_pcg_stream": (None, [u32, u32, vp, vp]),
            "orc_encode_normal": (u32, [vp]),
            "orc_decode_normal": (None, [u32, vp]),
            "orc_offset_ray_origin": (None, [vp, vp, vp]),
            "orc_f32_to_f16": (C.c_uint16, [f32]),
            "orc_frame": (None, [i32, vp, vp, f32, vp, vp]),
            "orc_equal_area_uv": (None, [vp, u64, vp]),
            "orc_surface_frame": (None, [vp, u32, vp, vp]),
            "orc_f16_to_f32": (f32, [C.c_uint16]),
            "orc_camera_look_at": (None, [vp, vp, vp, f32, f32, f32, f32, i32, u32, C.POINTER(abi.Camera)]),
            "orc_camera_rays": (None, [C.POINTER(abi.Camera), u32, u32, vp]),
            "orc_scene_create": (vp, [C.POINTER(abi.SceneDesc), i32]),
            "orc_scene_create_triangles": (vp, [vp, u32, vp, u32, i32, i32]),
            "orc_scene_destroy": (None, [vp]),
            "orc_trace_closest": (None, [vp, vp, u64, vp, i32]),
            "orc_gbuffer": (None, [vp, C.POINTER(abi.Camera), u32, u32, u32, vp, i32]),
            "orc_gbuffer_albedo": (None, [vp, C.POINTER(abi.Camera), u32, u32, vp, vp, i32]),
            "orc_pt_render": (u64, [vp, C.POINTER(abi.PtParams), C.POINTER(abi.Camera), u32, u32, u32, vp,
                                    u32, u32, u32, u32, vp, vp, i32]),
            "orc_bsdf": (None, [C.POINTER(BsdfCfg), i32, vp, vp, u32, u64, vp, vp, vp, vp]),
            "orc_math": (None, [i32, vp, vp, u64, vp]),
            "orc_guide_forward": (None, [C.POINTER(GuideSpec), vp, vp, vp, vp, u32, vp, vp, i32]),
            "orc_guide_gradient": (f32, [C.POINTER(GuideSpec), vp, vp, vp, u32, vp, vp, i32]),
            "orc_guide_sh": (None, [vp, vp]),
            "orc_guide_mixture_pdf": (C.c_double, [vp, vp, vp, vp, vp]),
            "orc_vmf_pdf": (C.c_double, [C.c_double, C.c_double]),
            "orc_mcpg_create": (vp, [C.POINTER(abi.McpgConfig)]),
            "orc_mcpg_destroy": (None, [vp]),
            "orc_mcpg_grids": (None, [vp, vp, vp, vp, vp]),
            "orc_mcpg_render": (u64, [vp, vp, C.POINTER(abi.PtParams), C.POINTER(abi.Camera), u32, u32, u32, vp,
                                      vp]),
            "orc_mcpg_probe": (None, [i32, vp, u64, vp, u32]),
            "orc_mcpg_cells": (None, [C.POINTER(abi.McpgGridConfig), vp, vp, vp, u64, vp]),
            "orc_mcpg_add_sample": (None, [vp, vp, f32, vp, f32, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def camera_look_at(spec, frame: int = 0) -> abi.Camera:
    cam = abi.Camera()
    f3 = C.c_float * 3
    lib().orc_camera_look_at(C.cast(f3(*spec.eye), C.c_void_p), C.cast(f3(*spec.target), C.c_void_p),
                             C.cast(f3(*spec.up), C.c_void_p), spec.vfov, spec.aspect, spec.near, spec.far,
                             spec.jitter, frame, C.byref(cam))
    return cam


def camera_rays(cam: abi.Camera, w: int, h: int) -> np.ndarray:
    out = np.zeros(w * h, dtype=abi.RAY_DTYPE)
    lib().orc_camera_rays(C.byref(cam), w, h, _p(out))
    return out


class OracleScene:
    def __init__(self, scene=None, *, triangles=None, two_sided=True, use_bvh=True):
        if scene is not None:
            self._desc = scene.desc()
            self.h = lib().orc_scene_create(C.byref(self._desc), int(use_bvh))
        else:
            pos, idx = triangles
            pos = np.ascontiguousarray(pos, dtype=np.float32)
            idx = np.ascontiguousarray(idx, dtype=np.uint32)
            self.h = lib().orc_scene_create_triangles(_p(pos), len(pos), _p(idx), len(idx) // 3,
                                                      int(two_sided), int(use_bvh))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def trace(self, rays: np.ndarray, threads: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), dtype=abi.HIT_DTYPE)
        lib().orc_trace_closest(self.h, _p(rays), len(rays), _p(hits), threads)
        return hits

    def gbuffer(self, cam, w, h, mask=0xFF, threads: int = 0) -> np.ndarray:
        out = np.zeros(w * h, dtype=abi.PACKED_HIT_DTYPE)
        lib().orc_gbuffer(self.h, C.byref(cam), w, h, mask, _p(out), threads)
        return out

    def gbuffer_albedo(self, cam, w, h, gbuffer, threads: int = 0) -> np.ndarray:
        """The GBuffer albedo texture, float32[h, w, 4] = (albedo.rgb, 1)."""
        out = np.zeros((h, w, 4), dtype=np.float32)
        gb = np.ascontiguousarray(gbuffer)
        lib().orc_gbuffer_albedo(self.h, C.byref(cam), w, h, _p(gb), _p(out), threads)
        return out

    def render(self, params: abi.PtParams, cam, frame, w, h, gbuffer, rect=None, threads: int = 0):
        """Returns (irradiance float32[ch, cw, 4], signature uint32[ch, cw], continuation segments)."""
        x0, y0, cw, ch = rect if rect else (0, 0, w, h)
        out = np.zeros((ch, cw, 4), dtype=np.float32)
        sig = np.zeros((ch, cw), dtype=np.uint32)
        gb = np.ascontiguousarray(gbuffer)
        segs = lib().orc_pt_render(self.h, C.byref(params), C.byref(cam), frame, w, h, _p(gb), x0, y0, cw, ch,
                                   _p(out), _p(sig), threads)
        return out, sig, int(segs)


MATH_OPS = {"sin": 0, "cos": 1, "atan": 2, "atan2": 3, "acos": 4, "log2": 5, "exp2": 6, "exp": 7, "pow": 8}


def math(op: str, x, y=None) -> np.ndarray:
    """The oracle's elementary functions (atan2(x, y) and pow(x, y) take two arguments)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    yy = None if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float32)
    out = np.zeros_like(x)
    lib().orc_math(MATH_OPS[op], _p(x), _p(yy), x.size, _p(out))
    return out


def frame(op: str, a, b=None, hs: float = 1.0, c=None) -> np.ndarray:
    """Shading frame rows (T, B, N) as a 3x3 array: op "normal" = make_frame(a), "nt" =
    OrthonormalFrame(a, b, hs), "normal_map" = OrthonormalFrame(a, b, hs).apply_normal_map(c)."""
    codes = {"normal": 0, "nt": 1, "normal_map": 2}
    f = lambda v: np.ascontiguousarray(np.zeros(3) if v is None else v, dtype=np.float32)
    a, b, c = f(a), f(b), f(c)
    out = np.zeros(9, np.float32)
    lib().orc_frame(codes[op], _p(a), _p(b), float(hs), _p(c), _p(out))
    return out.reshape(3, 3)


def equal_area_uv(dirs) -> np.ndarray:
    """EqualAreaOctMap.dir_to_uv over float32 directions (n, 3) -> (n, 2)."""
    d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
    uv = np.zeros((len(d), 2), np.float32)
    lib().orc_equal_area_uv(_p(d), len(d), _p(uv))
    return uv


def surface_frame(vertices: np.ndarray, flags: int, bary) -> dict:
    """Geometric part of Scene.get_surface for one triangle (3 abi.VERTEX_DTYPE records)."""
    v = np.ascontiguousarray(vertices)
    assert v.nbytes == 72
    b = np.ascontiguousarray(bary, dtype=np.float32)
    out = np.zeros(15, np.float32)
    lib().orc_surface_frame(_p(v), int(flags), _p(b), _p(out))
    o = out.reshape(5, 3)
    return {"T": o[0], "B": o[1], "N": o[2], "face_normal": o[3], "vertex_normal": o[4]}


def encode_normal(n) -> int:
    a = np.ascontiguousarray(n, dtype=np.float32)
    return int(lib().orc_encode_normal(_p(a)))


def decode_normal(e: int) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_decode_normal(int(e), _p(out))
    return out


def offset_ray_origin(p, n) -> np.ndarray:
    a = np.ascontiguousarray(p, dtype=np.float32)
    b = np.ascontiguousarray(n, dtype=np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_offset_ray_origin(_p(a), _p(b), _p(out))
    return out


def pcg4d16(x, y, z, w) -> int:
    return int(lib().orc_pcg4d16(x, y, z, w))


def pcg_stream(seed: int, n: int):
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    lib().orc_pcg_stream(seed, n, _p(u), _p(f))
    return u, f


def bsdf(cfg: BsdfCfg, op: str, wi: np.ndarray, wo: np.ndarray | None = None, seed: int = 1):
    wi = np.ascontiguousarray(wi, dtype=np.float32)
    n = len(wi)
    o_wo = np.zeros((n, 3), np.float32)
    o_pdf = np.zeros(n, np.float32)
    o_w = np.zeros((n, 3), np.float32)
    o_v = np.zeros(n, np.int32)
    wo_a = None if wo is None else np.ascontiguousarray(wo, dtype=np.float32)
    code = {"sample_eval": 0, "eval": 1, "pdf": 2, "albedo": 3}[op]
    lib().orc_bsdf(C.byref(cfg), code, _p(wi), _p(wo_a), seed, n, _p(o_wo), _p(o_pdf), _p(o_w), _p(o_v))
    if op == "sample_eval":
        return o_wo, o_pdf, o_w, o_v.astype(bool)
    if op in ("eval", "albedo"):
        return o_w
    return o_pdf


def bsdf_cfg(kind: int, albedo=(0.8, 0.8, 0.8), alpha=(0.3, 0.3), f0=0.04, eta=1.0 / 1.5, material=None,
             front_facing=True, ior=1.0) -> BsdfCfg:
    c = BsdfCfg()
    c.kind = kind
    for k in range(3):
        c.albedo[k] = albedo[k]
    c.alpha[0], c.alpha[1] = alpha
    c.f0, c.eta, c.front_facing, c.ior = f0, eta, int(front_facing), ior
    if material is not None:
        raw = np.ascontiguousarray(material).view(np.uint8).reshape(-1)[:128]
        C.memmove(c.mat, raw.ctypes.data, 128)
    return c


def pt_params(spp=4, max_path_length=4, emission_on_primary=True, mask=0xFF, mode=abi.MODE_BSDF,
              record_signature=True, demodulate_albedo=False) -> abi.PtParams:
    p = abi.PtParams()
    p.spp, p.max_path_length, p.emission_on_primary = spp, max_path_length, int(emission_on_primary)
    p.demodulate_albedo, p.instance_mask, p.mode = int(demodulate_albedo), mask, mode
    p.record_signature = int(record_signature)
    return p


# ------------------------------------------------------------------------------------------
# learned guide (oracle_guide.cpp)

def guide_spec(res, table: int, offset, scale) -> GuideSpec:
    g = GuideSpec()
    for k in range(8):
        g.res[k] = int(res[k])
    g.table = int(table)
    for a in range(3):
        g.offset[a], g.scale[a] = float(offset[a]), float(scale[a])
    return g


def guide_forward(spec: GuideSpec, weights: np.ndarray, grid: np.ndarray, pos: np.ndarray, wo: np.ndarray,
                  exact: bool = False, encoding: bool = False):
    """Network outputs [n, 40] (and the encodings [n, 64]); mirror rounding unless exact."""
    W = np.ascontiguousarray(weights, np.float32)
    G = np.ascontiguousarray(grid, np.float32)
    P = np.ascontiguousarray(pos, np.float32)
    D = np.ascontiguousarray(wo, np.float32)
    n = len(P)
    out = np.zeros((n, 40), np.float32)
    enc = np.zeros((n, 64), np.float32) if encoding else None
    lib().orc_guide_forward(C.byref(spec), _p(W), _p(G), _p(P), _p(D), n, _p(out), _p(enc), int(exact))
    return (out, enc) if encoding else out


def guide_gradient(spec: GuideSpec, weights: np.ndarray, grid: np.ndarray, records: np.ndarray, exact: bool = False):
    """(loss, dW [5, 64, 64], dgrid) of the training loss over records [n, 12]."""
    W = np.ascontiguousarray(weights, np.float32)
    G = np.ascontiguousarray(grid, np.float32)
    R = np.ascontiguousarray(records, np.float32)
    dW = np.zeros_like(W)
    dG = np.zeros_like(G)
    loss = lib().orc_guide_gradient(C.byref(spec), _p(W), _p(G), _p(R), len(R), _p(dW), _p(dG), int(exact))
    return loss, dW, dG


def guide_sh(d) -> np.ndarray:
    d = np.ascontiguousarray(d, np.float32)
    o = np.zeros(16, np.float32)
    lib().orc_guide_sh(_p(d), _p(o))
    return o


def guide_mixture(out40: np.ndarray, d) -> tuple[float, np.ndarray, np.ndarray, np.ndarray]:
    """(q(d), pi[8], kappa[8], mu[8, 3]) of one output vector."""
    o = np.ascontiguousarray(out40, np.float32)
    dd = np.ascontiguousarray(d, np.float32)
    pi, ka, mu = np.zeros(8, np.float32), np.zeros(8, np.float32), np.zeros((8, 3), np.float32)
    q = lib().orc_guide_mixture_pdf(_p(o), _p(dd), _p(pi), _p(ka), _p(mu))
    return q, pi, ka, mu


def vmf_pdf(kappa: float, cos_mu: float) -> float:
    return lib().orc_vmf_pdf(kappa, cos_mu)


# ---- Markov-chain path guiding (oracle_mcpg.h) ------------------------------------------
# orc_mcpg_cfg has the byte layout of mrt_mcpg_config (abi.McpgConfig).

class McpgOracle:
    """The MCPG renderer's state and frames on the CPU, one pixel and sample after another."""

    def __init__(self, cfg: abi.McpgConfig):
        self.cfg = cfg
        self.h = lib().orc_mcpg_create(C.byref(cfg))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mcpg_destroy(self.h)
            self.h = None

    def render(self, scene: OracleScene, params: abi.PtParams, cam, frame: int, w: int, h: int, gbuffer):
        out = np.zeros((h, w, 4), np.float32)
        gb = np.ascontiguousarray(gbuffer)
        segs = lib().orc_mcpg_render(scene.h, self.h, C.byref(params), C.byref(cam), frame, w, h, _p(gb), _p(out))
        return out, int(segs)

    def grids(self):
        n, m = self.cfg.mc.buffer_size, self.cfg.lc.buffer_size
        mk = np.zeros((n, 2), np.uint32)
        ms = np.zeros(n, abi.MCPG_STATE_DTYPE)
        lk = np.zeros((m, 2), np.uint32)
        le = np.zeros(m, abi.LC_ENTRY_DTYPE)
        lib().orc_mcpg_grids(self.h, _p(mk), _p(ms), _p(lk), _p(le))
        return mk, ms, lk, le


def vmf_pdf_tokuyoshi(kappa, mu, w) -> np.ndarray:
    """The oracle's vMF pdf (float32) for rows kappa[n], mu[n, 3], w[n, 3]."""
    a = np.zeros((len(w), 7), np.float32)
    a[:, 0], a[:, 1:4], a[:, 4:7] = kappa, mu, w
    out = np.zeros(len(w), np.float32)
    lib().orc_mcpg_probe(0, _p(a), len(w), _p(out), 0)
    return out


def vmf_sample_tokuyoshi(kappa, mu, seed: int = 1) -> np.ndarray:
    a = np.zeros((len(kappa), 4), np.float32)
    a[:, 0], a[:, 1:4] = kappa, mu
    out = np.zeros((len(kappa), 3), np.float32)
    lib().orc_mcpg_probe(1, _p(a), len(kappa), _p(out), seed)
    return out


def mcpg_scalar(op: str, x) -> np.ndarray:
    """log1p / expm1 / x_over_expm1 as the oracle evaluates them (float32)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    lib().orc_mcpg_probe({"log1p": 2, "expm1": 3, "x_over_expm1": 4}[op], _p(x), len(x), _p(out), 0)
    return out


def mcpg_cells(grid: abi.McpgGridConfig, pos, nrm, levels) -> np.ndarray:
    """(slot, key) of get_hash_grid_pos for each position / normal / level (exact cells)."""
    P = np.ascontiguousarray(pos, np.float32)
    N = np.ascontiguousarray(nrm, np.float32)
    L = np.ascontiguousarray(levels, np.uint32)
    out = np.zeros((len(P), 2), np.uint32)
    lib().orc_mcpg_cells(C.byref(grid), _p(P), _p(N), _p(L), len(P), _p(out))
    return out


def mcpg_add_sample(state: np.ndarray, pos, w: float, tgt, prior: float = 0.2):
    """Applies MCState.add_sample to a copy of a 32-byte state; returns (state, (kappa, mu))."""
    s = np.array(state, dtype=abi.MCPG_STATE_DTYPE).copy()
    p = np.ascontiguousarray(pos, np.float32)
    t = np.ascontiguousarray(tgt, np.float32)
    kv = np.zeros(4, np.float32)
    lib().orc_mcpg_add_sample(_p(s), _p(p), w, _p(t), prior, _p(kv))
    return s, kv
