"""Python host side over the C ABI (include/merian_amd.h): contexts, scenes, ray queries and
pt frames. Every call goes to libmerian_amd.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .scenes import CameraSpec, Scene


class MrtError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class DeviceScene:
    def __init__(self, ctx: "Context", handle: C.c_void_p, host: Scene | None):
        self.ctx, self.h, self.host = ctx, handle, host

    def info(self) -> dict:
        t, n, b = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx._check(self.ctx.lib.mrt_scene_info(self.h, C.byref(t), C.byref(n), C.byref(b)))
        return {"triangles": t.value, "nodes": n.value, "device_bytes": b.value}

    def destroy(self):
        if self.h:
            self.ctx.lib.mrt_scene_destroy(self.ctx.h, self.h)
            self.h = None


class Context:
    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = abi.load_library()
        h = C.c_void_p()
        st = self.lib.mrt_create(device, hbm_budget, C.byref(h))
        if st != abi.MRT_OK:
            raise MrtError(st, f"mrt_create(device={device}) failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.mrt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st: int):
        if st != abi.MRT_OK:
            raise MrtError(st, self.lib.mrt_last_error(self.h).decode())

    @property
    def memory_in_use(self) -> int:
        return int(self.lib.mrt_memory_in_use(self.h))

    def synchronize(self):
        self._check(self.lib.mrt_synchronize(self.h))

    # -- scenes -----------------------------------------------------------------------------
    def scene(self, s: Scene) -> DeviceScene:
        d = s.desc()
        h = C.c_void_p()
        self._check(self.lib.mrt_scene_create(self.h, C.byref(d), C.byref(h)))
        return DeviceScene(self, h, s)

    def triangles(self, positions: np.ndarray, indices: np.ndarray, two_sided: bool = True) -> DeviceScene:
        pos = np.ascontiguousarray(positions, dtype=np.float32)
        idx = np.ascontiguousarray(indices, dtype=np.uint32)
        h = C.c_void_p()
        self._check(self.lib.mrt_scene_create_triangles(self.h, _p(pos), len(pos), _p(idx), len(idx) // 3,
                                                        int(two_sided), C.byref(h)))
        return DeviceScene(self, h, None)

    # -- cameras ----------------------------------------------------------------------------
    def camera(self, spec: CameraSpec, frame: int = 0) -> abi.Camera:
        cam = abi.Camera()
        f3 = C.c_float * 3
        self._check(self.lib.mrt_camera_look_at(f3(*spec.eye), f3(*spec.target), f3(*spec.up), spec.vfov,
                                                spec.aspect, spec.near, spec.far, spec.jitter, frame,
                                                C.byref(cam)))
        return cam

    # -- queries ----------------------------------------------------------------------------
    def trace(self, scene: DeviceScene, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays)
        assert rays.dtype == abi.RAY_DTYPE
        hits = np.zeros(len(rays), dtype=abi.HIT_DTYPE)
        self._check(self.lib.mrt_trace_closest(self.h, scene.h, _p(rays), len(rays), _p(hits)))
        return hits

    def trace_dev(self, scene: DeviceScene, d_rays: int, n: int, d_hits: int, stream: int | None = None):
        self._check(self.lib.mrt_trace_closest_dev(self.h, scene.h, C.c_void_p(d_rays), n, C.c_void_p(d_hits),
                                                   C.c_void_p(stream) if stream else None))

    def camera_rays_dev(self, cam: abi.Camera, w: int, h: int, d_rays: int):
        self._check(self.lib.mrt_camera_rays_dev(self.h, C.byref(cam), w, h, C.c_void_p(d_rays), None))

    def gbuffer_dev(self, scene: DeviceScene, cam: abi.Camera, w: int, h: int, d_out: int, mask: int = 0xFF):
        self._check(self.lib.mrt_gbuffer_dev(self.h, scene.h, C.byref(cam), w, h, mask, C.c_void_p(d_out), None))

    def pt_render_dev(self, scene: DeviceScene, params: abi.PtParams, cam: abi.Camera, frame: int, w: int, h: int,
                      d_gbuffer: int, d_out: int, d_sig: int | None = None, stats: bool = False):
        st = abi.PtStats() if stats else None
        self._check(self.lib.mrt_pt_render_dev(self.h, scene.h, C.byref(params), C.byref(cam), frame, w, h,
                                               C.c_void_p(d_gbuffer), C.c_void_p(d_out),
                                               C.c_void_p(d_sig) if d_sig else None,
                                               C.byref(st) if st is not None else None, None))
        return st

    def render(self, scene: DeviceScene, params: abi.PtParams, cam: abi.Camera, frame: int, w: int, h: int,
               signature: bool = True, gbuffer: bool = True):
        """One frame (gbuffer + pt) with host results: (irradiance[h, w, 4], signature, gbuffer, stats)."""
        irr = np.zeros((h, w, 4), np.float32)
        sig = np.zeros((h, w), np.uint32) if signature else None
        gb = np.zeros(w * h, abi.PACKED_HIT_DTYPE) if gbuffer else None
        st = abi.PtStats()
        self._check(self.lib.mrt_render_frame(self.h, scene.h, C.byref(params), C.byref(cam), frame, w, h, _p(irr),
                                              _p(sig), _p(gb), C.byref(st)))
        return irr, sig, gb, st


def pt_params(spp=4, max_path_length=4, emission_on_primary=True, mask=0xFF, mode=abi.MODE_BSDF,
              record_signature=True) -> abi.PtParams:
    p = abi.PtParams()
    p.spp, p.max_path_length, p.emission_on_primary = spp, max_path_length, int(emission_on_primary)
    p.demodulate_albedo, p.instance_mask, p.mode, p.record_signature = 0, mask, mode, int(record_signature)
    return p
