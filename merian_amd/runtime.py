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


HIP_STREAM_LEGACY = 1  # hipStreamLegacy: the HIP null stream


def _stream(stream: int | None):
    """The stream argument of a _dev entry point. None selects the context's own stream; a HIP
    stream handle selects that stream, where handle 0 (the null stream, torch's default stream)
    is passed as hipStreamLegacy, because a NULL argument would select the context's stream."""
    if stream is None:
        return None
    return C.c_void_p(HIP_STREAM_LEGACY if int(stream) == 0 else int(stream))


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

    def tuning(self) -> abi.Tuning:
        t = abi.Tuning()
        self.ctx._check(self.ctx.lib.mrt_scene_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, t: abi.Tuning):
        self.ctx._check(self.ctx.lib.mrt_scene_set_tuning(self.ctx.h, self.h, C.byref(t)))

    def tune(self, cam: abi.Camera, w: int, h: int, params: abi.PtParams, budget_ms: float = 20000.0) -> abi.Tuning:
        """Autotunes this scene's BVH and launch settings on the given view (mrt_tune)."""
        r = abi.TuneRequest()
        r.camera, r.width, r.height, r.params, r.time_budget_ms = cam, w, h, params, budget_ms
        t = abi.Tuning()
        self.ctx._check(self.ctx.lib.mrt_tune(self.ctx.h, self.h, C.byref(r), C.byref(t)))
        return t


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
                                                   _stream(stream)))

    def trace_noalloc_dev(self, scene: DeviceScene, d_rays: int, n: int, d_hits: int, stream: int | None = None):
        """mrt_trace_closest_noalloc_dev: as trace_dev, with the library's fixed scratch (no allocation)."""
        self._check(self.lib.mrt_trace_closest_noalloc_dev(self.h, scene.h, C.c_void_p(d_rays), n,
                                                           C.c_void_p(d_hits), _stream(stream)))

    def trace_count_dev(self, scene: DeviceScene, d_rays: int, n: int, d_hits: int) -> tuple[int, int]:
        """As trace_dev with the counting kernel (synchronous): (node visits, triangle tests)."""
        counts = (C.c_uint64 * 2)()
        self._check(self.lib.mrt_trace_count_dev(self.h, scene.h, C.c_void_p(d_rays), n, C.c_void_p(d_hits), counts))
        return int(counts[0]), int(counts[1])

    def camera_rays_dev(self, cam: abi.Camera, w: int, h: int, d_rays: int, stream: int | None = None):
        self._check(self.lib.mrt_camera_rays_dev(self.h, C.byref(cam), w, h, C.c_void_p(d_rays), _stream(stream)))

    def gbuffer_dev(self, scene: DeviceScene, cam: abi.Camera, w: int, h: int, d_out: int, mask: int = 0xFF,
                    stream: int | None = None):
        self._check(self.lib.mrt_gbuffer_dev(self.h, scene.h, C.byref(cam), w, h, mask, C.c_void_p(d_out),
                                             _stream(stream)))

    def gbuffer_albedo_dev(self, scene: DeviceScene, cam: abi.Camera, w: int, h: int, d_gbuffer: int, d_albedo: int,
                           stream: int | None = None):
        """mrt_gbuffer_albedo_dev: the GBuffer albedo texture (float4 per pixel) of a hit texture."""
        self._check(self.lib.mrt_gbuffer_albedo_dev(self.h, scene.h, C.byref(cam), w, h, C.c_void_p(d_gbuffer),
                                                    C.c_void_p(d_albedo), _stream(stream)))

    def pt_render_dev(self, scene: DeviceScene, params: abi.PtParams, cam: abi.Camera, frame: int, w: int, h: int,
                      d_gbuffer: int, d_out: int, d_sig: int | None = None, stats: bool = False,
                      stream: int | None = None):
        """Enqueues a frame on `stream`: a HIP stream handle such as
        torch.cuda.current_stream().cuda_stream (0, torch's default stream, is the HIP null
        stream), or None for the context's own stream, which does not wait for torch's."""
        st = abi.PtStats() if stats else None
        self._check(self.lib.mrt_pt_render_dev(self.h, scene.h, C.byref(params), C.byref(cam), frame, w, h,
                                               C.c_void_p(d_gbuffer), C.c_void_p(d_out),
                                               C.c_void_p(d_sig) if d_sig else None,
                                               C.byref(st) if st is not None else None,
                                               _stream(stream)))
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
              record_signature=True, demodulate_albedo=False) -> abi.PtParams:
    p = abi.PtParams()
    p.spp, p.max_path_length, p.emission_on_primary = spp, max_path_length, int(emission_on_primary)
    p.demodulate_albedo, p.instance_mask, p.mode = int(demodulate_albedo), mask, mode
    p.record_signature = int(record_signature)
    return p


class Job:
    """A submitted render job; keeps its host output buffer alive until it is done."""

    def __init__(self, server: "Server", job_id: int, irradiance: np.ndarray, desc: abi.JobDesc):
        self.server, self.id, self.irradiance, self.desc = server, job_id, irradiance, desc

    def poll(self) -> abi.JobStatus:
        st = abi.JobStatus()
        self.server._check(self.server.lib.mrt_job_poll(self.server.h, self.id, C.byref(st)))
        return st

    def wait(self, timeout_ms: int = -1) -> abi.JobStatus:
        st = abi.JobStatus()
        r = self.server.lib.mrt_job_wait(self.server.h, self.id, timeout_ms, C.byref(st))
        if r not in (abi.MRT_OK, abi.MRT_ERR_NOT_READY):
            self.server._check(r)
        return st

    def cancel(self) -> bool:
        return self.server.lib.mrt_job_cancel(self.server.h, self.id) == abi.MRT_OK


class Server:
    """Continuous-batching render server over mrt_server_* (one scheduler thread in the library)."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, max_slots: int = 4, max_paths_per_job: int = 0,
                 use_graphs: bool = True):
        self.ctx, self.lib = ctx, ctx.lib
        cfg = abi.ServerConfig(hbm_budget, max_slots, max_paths_per_job, int(use_graphs), 0)
        h = C.c_void_p()
        ctx._check(self.lib.mrt_server_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def _check(self, st: int):
        if st != abi.MRT_OK:
            raise MrtError(st, self.lib.mrt_server_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.mrt_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def desc(scene: DeviceScene, params: abi.PtParams, cam: abi.Camera, frame: int, w: int, h: int,
             out: np.ndarray | None = None) -> abi.JobDesc:
        d = abi.JobDesc()
        d.scene, d.camera, d.params, d.width, d.height, d.frame = scene.h, cam, params, w, h, frame
        d.irradiance = out.ctypes.data if out is not None else None
        return d

    def job_bytes(self, scene, params, cam, frame, w, h) -> int:
        out = np.zeros(4, np.float32)
        d = self.desc(scene, params, cam, frame, w, h, out)
        b = C.c_uint64()
        self._check(self.lib.mrt_job_bytes(self.h, C.byref(d), C.byref(b)))
        return b.value

    def submit(self, scene: DeviceScene, params: abi.PtParams, cam: abi.Camera, frame: int, w: int, h: int) -> Job:
        out = np.zeros((h, w, 4), np.float32)
        d = self.desc(scene, params, cam, frame, w, h, out)
        jid = C.c_uint64()
        self._check(self.lib.mrt_job_submit(self.h, C.byref(d), C.byref(jid)))
        return Job(self, jid.value, out, d)

    def stats(self) -> abi.ServerStats:
        s = abi.ServerStats()
        self._check(self.lib.mrt_server_stats_get(self.h, C.byref(s)))
        return s

    def reset_stats(self):
        self._check(self.lib.mrt_server_stats_reset(self.h))


def guide_config(seed: int = 1, log2_table: int = 0, base_resolution: int = 0, max_resolution: int = 0,
                 train_batch: int = 0, train_steps: int = 0, max_records: int = 0, learning_rate: float = 0.0,
                 guide_fraction: float = 0.0, bbox=None, guided_bounces: int = 0) -> abi.GuideConfig:
    c = abi.GuideConfig()
    c.seed, c.log2_table, c.base_resolution, c.max_resolution = seed, log2_table, base_resolution, max_resolution
    c.train_batch, c.train_steps, c.max_records = train_batch, train_steps, max_records
    c.learning_rate, c.guide_fraction = learning_rate, guide_fraction
    c.guided_bounces = guided_bounces
    if bbox is not None:
        for k in range(3):
            c.bbox_min[k], c.bbox_max[k] = bbox[0][k], bbox[1][k]
    return c


def guide_layout(cfg: abi.GuideConfig):
    """(per-level resolutions, table entries) of a guide configuration (host computation)."""
    lib = abi.load_library()
    res = (C.c_uint32 * abi.GUIDE_LEVELS)()
    t = C.c_uint32()
    st = lib.mrt_guide_layout(C.byref(cfg), res, C.byref(t))
    if st != abi.MRT_OK:
        raise MrtError(st, "mrt_guide_layout: configuration out of range")
    return list(res), t.value


class Guide:
    """Learned guide over mrt_guide_* (network parameters and training state on the device)."""

    def __init__(self, ctx: Context, scene: DeviceScene, cfg: abi.GuideConfig | None = None):
        self.ctx, self.lib = ctx, ctx.lib
        self.cfg = cfg if cfg is not None else guide_config()
        h = C.c_void_p()
        ctx._check(self.lib.mrt_guide_create(ctx.h, scene.h, C.byref(self.cfg), C.byref(h)))
        self.h = h
        self.resolutions, self.table = guide_layout(self.cfg)

    def close(self):
        if self.h:
            self.lib.mrt_guide_destroy(self.ctx.h, self.h)
            self.h = None

    def normalisation(self):
        o, s = (C.c_float * 3)(), (C.c_float * 3)()
        self.ctx._check(self.lib.mrt_guide_normalisation(self.h, o, s))
        return list(o), list(s)

    def params(self):
        W = np.zeros((abi.GUIDE_LAYERS, 64, 64), np.float32)
        G = np.zeros((abi.GUIDE_LEVELS, self.table, abi.GUIDE_FEATURES), np.float32)
        self.ctx._check(self.lib.mrt_guide_get_params(self.ctx.h, self.h, _p(W), _p(G)))
        return W, G

    def set_params(self, W: np.ndarray | None, G: np.ndarray | None):
        W = None if W is None else np.ascontiguousarray(W, np.float32)
        G = None if G is None else np.ascontiguousarray(G, np.float32)
        self.ctx._check(self.lib.mrt_guide_set_params(self.ctx.h, self.h, _p(W), _p(G)))

    def forward(self, pos: np.ndarray, wo: np.ndarray) -> np.ndarray:
        P = np.ascontiguousarray(pos, np.float32)
        D = np.ascontiguousarray(wo, np.float32)
        out = np.zeros((len(P), abi.GUIDE_OUTPUTS), np.float32)
        self.ctx._check(self.lib.mrt_guide_forward(self.ctx.h, self.h, _p(P), _p(D), len(P), _p(out)))
        return out

    def gradient(self, records: np.ndarray):
        R = np.ascontiguousarray(records, np.float32)
        dW = np.zeros((abi.GUIDE_LAYERS, 64, 64), np.float32)
        dG = np.zeros((abi.GUIDE_LEVELS, self.table, abi.GUIDE_FEATURES), np.float32)
        loss = C.c_float()
        self.ctx._check(self.lib.mrt_guide_gradient(self.ctx.h, self.h, _p(R), len(R), _p(dW), _p(dG),
                                                    C.byref(loss)))
        return loss.value, dW, dG

    def train(self, records: np.ndarray, steps: int) -> float:
        R = np.ascontiguousarray(records, np.float32)
        loss = C.c_float()
        self.ctx._check(self.lib.mrt_guide_train(self.ctx.h, self.h, _p(R), len(R), steps, C.byref(loss)))
        return loss.value

    def render_dev(self, scene: DeviceScene, params: abi.PtParams, cam: abi.Camera, frame: int, w: int, h: int,
                   d_gbuffer: int, d_out: int, train: bool = True, stats: bool = False, stream: int | None = None):
        st = abi.PtStats() if stats else None
        self.ctx._check(self.lib.mrt_guided_render_dev(self.ctx.h, scene.h, self.h, C.byref(params), C.byref(cam),
                                                       frame, w, h, C.c_void_p(d_gbuffer), C.c_void_p(d_out),
                                                       int(train), C.byref(st) if st is not None else None,
                                                       _stream(stream)))
        return st


def mcpg_config(**kw) -> abi.McpgConfig:
    """The MCPG node's defaults (mrt_mcpg_default_config) with overrides: top-level fields by
    name, grid fields as mc_<field> / lc_<field> (e.g. mc_buffer_size=4096)."""
    c = abi.McpgConfig()
    abi.load_library().mrt_mcpg_default_config(C.byref(c))
    for k, v in kw.items():
        if k.startswith("mc_") and hasattr(c.mc, k[3:]):
            setattr(c.mc, k[3:], v)
        elif k.startswith("lc_") and hasattr(c.lc, k[3:]):
            setattr(c.lc, k[3:], v)
        elif hasattr(c, k):
            setattr(c, k, v)
        else:
            raise KeyError(k)
    return c


class Mcpg:
    """Markov-chain path guiding state over mrt_mcpg_* (the two hash grids on the device)."""

    def __init__(self, ctx: Context, cfg: abi.McpgConfig | None = None):
        self.ctx, self.lib = ctx, ctx.lib
        self.cfg = cfg if cfg is not None else mcpg_config()
        h = C.c_void_p()
        ctx._check(self.lib.mrt_mcpg_create(ctx.h, C.byref(self.cfg), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.mrt_mcpg_destroy(self.ctx.h, self.h)
            self.h = None

    def reset(self):
        self.ctx._check(self.lib.mrt_mcpg_reset(self.ctx.h, self.h))

    def render_dev(self, scene: DeviceScene, params: abi.PtParams, cam: abi.Camera, frame: int, w: int, h: int,
                   d_gbuffer: int, d_out: int, stats: bool = False, stream: int | None = None):
        st = abi.PtStats() if stats else None
        self.ctx._check(self.lib.mrt_mcpg_render_dev(self.ctx.h, scene.h, self.h, C.byref(params), C.byref(cam),
                                                     frame, w, h, C.c_void_p(d_gbuffer), C.c_void_p(d_out),
                                                     C.byref(st) if st is not None else None,
                                                     _stream(stream)))
        return st

    def grids(self):
        """(mc keys uint32[n, 2], mc states MCPG_STATE_DTYPE[n], lc keys uint32[m, 2], lc entries LC_ENTRY_DTYPE[m])."""
        n, m = self.cfg.mc.buffer_size, self.cfg.lc.buffer_size
        mk = np.zeros((n, 2), np.uint32)
        ms = np.zeros(n, abi.MCPG_STATE_DTYPE)
        lk = np.zeros((m, 2), np.uint32)
        le = np.zeros(m, abi.LC_ENTRY_DTYPE)
        self.ctx._check(self.lib.mrt_mcpg_read_grids(self.ctx.h, self.h, _p(mk), _p(ms), _p(lk), _p(le)))
        return mk, ms, lk, le
