#!/usr/bin/env python
"""Benchmark: path segments per second of one pt frame on one MI355X.

Workload (the headline metric's configuration): 1920x1080, 64 spp, max path length 8, the seeded
procedural scene of 1,999,880 triangles (256 displaced tori instances + an open room), 256
materials (Lambert / GGX conductor / dielectric 1.5), 2048x1024 lat-long sky with a sun of
5e4:1 dynamic range. A step = one frame: GBuffer pass + pt (BSDF-only estimator, the
reference's). Inputs (scene, BVH, buffers) are resident in HBM before timing starts.

Further legs in the same line: `guided` (equal-time relMSE of the learned guide and MCPG against
plain pt), `serving` (Poisson job mix, p50/p99), `ray_query` (1M random triangles, 16M
incoherent + 4096^2 coherent closest-hit rays, Mrays/s), `small` (the small plain-pt
configuration: 81,922 triangles, 512^2, 4 spp, L = 4) and `cpu_baseline` (the CPU oracle on
bounded samples of the frame and of the ray queries, and on the whole small frame).

Prints one JSON line. --gpus N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpath-segments/s at 1080p; equal-time relMSE guided vs plain; p99 job latency"
HBM_PEAK_GBS = 8000.0
NODE_BYTES = {4: 128, 8: 80}  # BVH4Node (fp32 boxes), BVH8Node (8-bit quantized boxes)
TRI_BYTES = 48        # TriRecord
RAY_IO_BYTES = 32 + 16  # queued ray in, hit out
# HBM bytes per dispatch of the intersect kernel, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes over this workload and calibrated against a plain copy (python -m merian_amd.profiling traffic)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_trace_queue_full_1080p.traffic.json")
# autotune result of the workload, applied before timing (re-run the search with --retune)
TUNING_CACHE = os.path.join(ROOT, "merian_amd", "tuning", "full_1080p.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, spec, frame, target_s: float, threads: int):
    """The CPU oracle (a port of the reference pt) on a bounded crop of the same frame."""
    from oracle import oracle as O
    W, H = spec.width, spec.height
    osc = O.OracleScene(scene)
    cam = O.camera_look_at(spec, frame)
    t0 = time.perf_counter()
    gb = osc.gbuffer(cam, W, H, threads=threads)
    t_gb = time.perf_counter() - t0
    p = O.pt_params(64, 8, record_signature=False)
    cw, ch = 32, 18
    # grow the centre crop until the timed render takes at least half of target_s
    t_pt = 0.0
    for _ in range(4):
        t0 = time.perf_counter()
        _, _, segs = osc.render(p, cam, frame, W, H, gb, rect=((W - cw) // 2, (H - ch) // 2, cw, ch),
                                threads=threads)
        t_pt = time.perf_counter() - t0
        if t_pt >= 0.5 * target_s or (cw == W and ch == H):
            break
        scale = min(4.0, max(1.2, (target_s / max(t_pt, 1e-3)) ** 0.5))
        cw, ch = min(W, int(cw * scale) // 2 * 2), min(H, int(ch * scale) // 2 * 2)
    npx = cw * ch
    t = t_pt + t_gb * npx / (W * H)
    return {"value": round((segs + npx) / t / 1e6, 4), "unit": "Mpath-segments/s", "cores": threads,
            "kind": "port",
            "sample": f"{cw}x{ch} centre crop of the same 1920x1080 frame, 64 spp, L=8 "
                      f"({segs + npx} segments, {t:.1f} s incl. the crop's share of the CPU gbuffer pass)"}


def cpu_small(threads: int, min_s: float = 3.0) -> dict:
    """The CPU oracle on whole small plain-pt frames (512^2, 4 spp, L = 4), gbuffer included,
    frames 0, 1, ... until at least `min_s` seconds have passed."""
    from merian_amd import scenes
    from oracle import oracle as O
    sc = scenes.plain_small(seed=0)
    spec = sc.camera
    osc = O.OracleScene(sc)
    p = O.pt_params(4, 4, record_signature=False)
    segs, f, t = 0, 0, 0.0
    while t < min_s:
        cam = O.camera_look_at(spec, f)
        t0 = time.perf_counter()
        gb = osc.gbuffer(cam, spec.width, spec.height, threads=threads)
        _, _, s_ = osc.render(p, cam, f, spec.width, spec.height, gb, threads=threads)
        t += time.perf_counter() - t0
        segs += s_ + spec.width * spec.height
        f += 1
    return {"value": round(segs / t / 1e6, 3), "unit": "Mpath-segments/s", "cores": threads, "kind": "port",
            "sample": f"{f} whole frames ({segs} segments, {t:.2f} s)"}


def small_leg(ctx, frames: int = 64) -> dict:
    """The small plain-pt configuration: 8x8 icospheres on a ground plane (81,922 triangles),
    constant white environment, 512x512, 4 spp, max path length 4; a step = GBuffer + pt frame
    with its statistics read back. The scene is autotuned on the view first (mrt_tune)."""
    import torch
    from merian_amd import scenes
    from merian_amd.runtime import pt_params
    sc = scenes.plain_small(seed=0)
    spec = sc.camera
    W, H = spec.width, spec.height
    ds = ctx.scene(sc)
    gb = torch.empty(W * H * 16, dtype=torch.uint8, device="cuda")
    out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    params = pt_params(spp=4, max_path_length=4, record_signature=False)
    t0 = time.perf_counter()
    ds.tune(ctx.camera(spec, 0), W, H, params, budget_ms=20000.0)  # mrt_tune on this view, untimed
    tuning = {k: v for k, v in ds.tuning().as_dict().items() if k not in ("ms_per_frame", "_reserved")}
    tune_s = time.perf_counter() - t0

    def frame(f):
        cam = ctx.camera(spec, f)
        ctx.gbuffer_dev(ds, cam, W, H, gb.data_ptr())
        return ctx.pt_render_dev(ds, params, cam, f, W, H, gb.data_ptr(), out.data_ptr(), stats=True)

    frame(0)
    torch.cuda.synchronize()
    segs = 0
    t0 = time.perf_counter()
    for f in range(1, frames + 1):
        st = frame(f)
        segs += st.primary_segments + st.continuation_segments
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ds.destroy()
    res = {"workload": f"{sc.triangle_count} triangles, 512x512, 4 spp, max path length 4, constant environment",
           "mseg_per_s": round(segs / dt / 1e6, 1), "ms_per_frame": round(dt * 1e3 / frames, 3),
           "segments_per_frame": segs // frames, "frames": frames, "tuning": tuning,
           "tuning_source": f"mrt_tune before timing ({tune_s:.1f} s)"}
    log(f"small leg: {res}")
    return res


def cpu_ray_query(threads: int) -> dict:
    """The CPU oracle's BVH traversal on bounded samples of the ray-query workload: the same 1M
    triangles, 1,048,576 incoherent rays of the same distribution and every 16th of the 4096^2
    coherent rays."""
    import numpy as np
    from merian_amd import scenes
    from oracle import oracle as O
    pos, idx = scenes.ray_query(seed=1, n_tri=1_000_000)
    osc = O.OracleScene(triangles=(pos, idx), use_bvh=True)
    spec = scenes.ray_query_camera(4096)
    sets = (("incoherent", scenes.incoherent_rays(1 << 20, seed=12)),
            ("coherent", np.ascontiguousarray(O.camera_rays(O.camera_look_at(spec, 0), 4096, 4096)[::16])))
    out = {"unit": "Mrays/s", "cores": threads, "kind": "port"}
    for name, rays in sets:
        t0 = time.perf_counter()
        osc.trace(rays, threads=threads)
        out[name] = round(len(rays) / (time.perf_counter() - t0) / 1e6, 3)
    out["sample"] = "1,048,576 incoherent rays and every 16th of the 4096x4096 coherent rays over the same 1M triangles"
    return out


def guided_leg(ctx, scene, L: int) -> dict:
    """Equal-time relMSE of guided vs plain pt on a 480x270 view of the 1080p scene (same camera,
    lower resolution), against a 16384-spp plain render of that view. The guided side trains
    online during its frames; both sides get the same wall time (timed around whole frames).
    The Markov-chain guiding (MCPG, the non-learned comparator) is measured the same way, with
    the MCPG example graph's grid settings, learning online from an empty state."""
    import numpy as np
    import torch
    from merian_amd import abi, scenes
    from merian_amd.runtime import Guide, Mcpg, guide_config, mcpg_config, pt_params
    w, h = 480, 270
    spec = scenes.CameraSpec(scene.camera.eye, scene.camera.target, scene.camera.up, scene.camera.vfov, w, h)
    ds = ctx.scene(scene)
    cam = ctx.camera(spec, 0)
    gb = torch.empty(w * h * 16, dtype=torch.uint8, device="cuda")
    out = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
    acc = torch.zeros(w * h * 4, dtype=torch.float64, device="cuda")
    ctx.gbuffer_dev(ds, cam, w, h, gb.data_ptr())

    def run(frames, params, guide=None, train=True, base=0, mcpg=None, keep=True):
        """Renders `frames` frames on the library's context stream; the device is synchronised
        after each frame before torch (another stream) accumulates it. Returns the mean image
        (None with keep=False: the frames' samples are discarded) and the wall time."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        acc.zero_()
        for f in range(frames):
            if mcpg is not None:
                mcpg.render_dev(ds, params, cam, base + f, w, h, gb.data_ptr(), out.data_ptr())
            elif guide is None:
                ctx.pt_render_dev(ds, params, cam, base + f, w, h, gb.data_ptr(), out.data_ptr())
            else:
                guide.render_dev(ds, params, cam, base + f, w, h, gb.data_ptr(), out.data_ptr(), train=train)
            torch.cuda.synchronize()
            if keep:
                acc.add_(out)
        torch.cuda.synchronize()
        img = (acc / frames).cpu().numpy().reshape(h, w, 4) if keep else None
        return img, time.perf_counter() - t0

    plain = pt_params(spp=64, max_path_length=L, record_signature=False)
    ref, t_ref = run(256, plain, base=100000)  # 16384 spp
    _, t8 = run(8, plain, base=300000, keep=False)
    ms_plain = 1e3 * t8 / 8
    # guided: TRAIN online-training frames (timed, samples discarded), then RENDER frames with
    # the guide frozen; plain gets the same total wall time
    TRAIN, RENDER = 16, 96
    g = Guide(ctx, ds, guide_config(seed=1, train_steps=16, guided_bounces=3))
    gp = pt_params(spp=64, max_path_length=L, mode=abi.MODE_GUIDED, record_signature=False)
    _, t_train = run(TRAIN, gp, guide=g, train=True, keep=False)
    gimg, t_render = run(RENDER, gp, guide=g, train=False, base=1000)
    g.close()
    t_g = t_train + t_render
    frames_p = max(1, int(round(t_g / (t8 / 8))))
    pimg, t_p = run(frames_p, plain)
    # MCPG: learning frames (timed, samples discarded), then frames that keep learning, as the
    # node runs; the MCPG example graph's grid settings
    m = Mcpg(ctx, mcpg_config(mc_tan_alpha_half=0.003, mc_distribution_dimension=1.5, lc_tan_alpha_half=0.012,
                              lc_buffer_size=8000009, lc_probe_count=2))
    _, tm_learn = run(4, plain, mcpg=m, base=5000, keep=False)
    mimg, tm_render = run(8, plain, mcpg=m, base=6000)
    m.close()
    t_m = tm_learn + tm_render
    frames_pm = max(1, int(round(t_m / (t8 / 8))))
    pmimg, t_pm = run(frames_pm, plain, base=20000)
    ds.destroy()
    rel = lambda x: float(np.mean((x[..., :3] - ref[..., :3]) ** 2 / (ref[..., :3] ** 2 + 1e-2)))
    res = {"workload": f"{w}x{h} view of the 1080p scene, L={L}, 64 spp per frame; reference 16384 spp plain",
           "protocol": f"guided (vertices 0-2): {TRAIN} online-training frames (train_steps 16, samples discarded) "
                       f"then {RENDER} frames with the guide frozen, all timed; plain: every frame of the same wall time",
           "relmse_guided": rel(gimg), "relmse_plain": rel(pimg), "seconds_guided": round(t_g, 3),
           "seconds_guided_training": round(t_train, 3), "seconds_plain": round(t_p, 3),
           "ms_per_frame_guided": round(1e3 * t_render / RENDER, 2), "ms_per_frame_plain": round(ms_plain, 2),
           "frames_guided": RENDER, "frames_plain": frames_p,
           "relmse_mcpg": rel(mimg), "seconds_mcpg": round(t_m, 3), "frames_mcpg": 8,
           "relmse_plain_at_mcpg_time": rel(pmimg), "frames_plain_at_mcpg_time": frames_pm,
           "mcpg_note": "MCPG guides rough glass: the GGX BTDF evaluation it shares with the plain pt is non-zero "
                        "where its pdf is 0, so guided refractions there carry unbounded weights (fireflies)",
           "reference_seconds": round(t_ref, 2)}
    log(f"guided leg: {res}")
    return res


def serving_leg(ctx, tune: bool = True) -> dict:
    """Serving mix: Poisson arrivals over 4 resident scenes (82k..1.0M triangles), resolutions
    {256^2, 512^2, 1024x768, 1920x1080}, spp {1, 4, 16, 64}, L in 4..8; arrival rate 80 % of the
    capacity measured with a closed batch. Latency = submit -> result in the host buffer."""
    import numpy as np
    from merian_amd import scenes
    from merian_amd.runtime import Server, pt_params
    rng = np.random.default_rng(3)
    resident = [scenes.plain_small(seed=101), scenes.full_scene(seed=102, instances=26, env_size=(512, 256)),
                scenes.plain_small(seed=105, subdiv=4), scenes.full_scene(seed=103, instances=128, env_size=(512, 256))]
    dss = [ctx.scene(s) for s in resident]
    res_choices = [(256, 256), (512, 512), (1024, 768), (1920, 1080)]
    tune_s = 0.0
    if tune:
        # each resident scene is autotuned once, on a 1024x768 view at 16 spp, L = 6 (untimed)
        t0 = time.perf_counter()
        for sc, d in zip(resident, dss):
            spec = scenes.CameraSpec(sc.camera.eye, sc.camera.target, sc.camera.up, sc.camera.vfov, 1024, 768)
            d.tune(ctx.camera(spec, 0), 1024, 768, pt_params(spp=16, max_path_length=6, record_signature=False),
                   budget_ms=5000.0)
        tune_s = time.perf_counter() - t0

    def job():
        k = int(rng.integers(len(dss)))
        w, h = res_choices[int(rng.integers(4))]
        spp = [1, 4, 16, 64][int(rng.integers(4))]
        L = int(rng.integers(4, 9))
        sc = resident[k]
        spec = scenes.CameraSpec(sc.camera.eye, sc.camera.target, sc.camera.up, sc.camera.vfov, w, h)
        f = int(rng.integers(1 << 16))
        return dss[k], pt_params(spp=spp, max_path_length=L, record_signature=False), ctx.camera(spec, f), f, w, h

    with Server(ctx, max_slots=4) as srv:
        batch = [job() for _ in range(16)]
        t0 = time.perf_counter()
        subs = [srv.submit(*j) for j in batch]
        for s_ in subs:
            s_.wait()
        capacity = len(batch) / (time.perf_counter() - t0)
        srv.reset_stats()
        lam = 0.8 * capacity
        jobs = [job() for _ in range(48)]
        arrivals = np.cumsum(rng.exponential(1.0 / lam, len(jobs)))
        subs = []
        t0 = time.perf_counter()
        for a, j in zip(arrivals, jobs):
            dt = a - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            subs.append(srv.submit(*j))
        for s_ in subs:
            s_.wait()
        wall = time.perf_counter() - t0
        st = srv.stats()
    for d in dss:
        d.destroy()
    res = {"workload": "48 Poisson-arrival jobs over 4 resident scenes (81,922 / 203,120 / 327,682 / "
                       "999,944 triangles), res {256^2, 512^2, 1024x768, 1920x1080}, spp {1,4,16,64}, L 4..8",
           "arrival_rate_per_s": round(lam, 2), "capacity_jobs_per_s": round(capacity, 2),
           "p50_ms": round(st.p50_ms, 2), "p99_ms": round(st.p99_ms, 2), "max_ms": round(st.max_ms, 2),
           "jobs_per_s": round(st.completed / wall, 2), "mseg_per_s": round(st.segments / wall / 1e6, 1),
           "peak_in_flight": st.peak_in_flight, "completed": st.completed,
           "tuning_source": f"mrt_tune per scene before the mix ({tune_s:.1f} s)" if tune else "defaults"}
    log(f"serving leg: {res}")
    return res


# Traversal settings of the ray-query leg, per ray set: the fastest of a grid search over BVH
# width {4, 8}, leaf size {1, 2, 4}, SAH node cost {0.5, 1, 2} (BVH4 leaf 2 only), refill lanes
# {8, 16, 24, 32, 48}, leaf-phase lanes {4, 8, 16, 32} and grid {50, 75, 100} % on this workload
# (re-run with --retune-rays). No setting changes a hit.
RAY_QUERY_TUNING = {
    "incoherent": dict(bvh_width=8, bvh_max_leaf=2, bvh_traversal_cost=1.0, refill_lanes=16, leaf_lanes=4,
                       grid_percent=100),
    "coherent": dict(bvh_width=4, bvh_max_leaf=2, bvh_traversal_cost=1.0, refill_lanes=48, leaf_lanes=8,
                     grid_percent=100),
}


def _apply(ds, settings: dict):
    t = ds.tuning()
    for k, v in settings.items():
        setattr(t, k, v)
    ds.set_tuning(t)


def _time_rays(ctx, ds, d_r, n, d_hits, stream, reps):
    """Average ms of `reps` closest-hit launches after one warm-up (HIP events on the launching stream)."""
    import torch
    with torch.cuda.stream(stream):
        ctx.trace_dev(ds, d_r.data_ptr(), n, d_hits.data_ptr(), stream=stream.cuda_stream)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        for _ in range(reps):
            ctx.trace_dev(ds, d_r.data_ptr(), n, d_hits.data_ptr(), stream=stream.cuda_stream)
        ev[1].record(stream)
    stream.synchronize()
    return ev[0].elapsed_time(ev[1]) / reps


def _search_ray_tuning(ctx, ds, d_r, n, d_hits, stream) -> dict:
    import itertools
    best = None
    trees = [(4, 2, 1.0), (4, 1, 1.0), (4, 4, 1.0), (8, 2, 1.0), (8, 4, 1.0), (4, 2, 2.0), (4, 2, 0.5)]
    for width, leaf, cost in trees:
        _apply(ds, dict(bvh_width=width, bvh_max_leaf=leaf, bvh_traversal_cost=cost))
        for refill, ll, grid in itertools.product([8, 16, 24, 32, 48], [4, 8, 16, 32], [50, 75, 100]):
            cand = dict(bvh_width=width, bvh_max_leaf=leaf, bvh_traversal_cost=cost, refill_lanes=refill,
                        leaf_lanes=ll, grid_percent=grid)
            _apply(ds, cand)
            ms = _time_rays(ctx, ds, d_r, n, d_hits, stream, 2)
            log(f"  {cand}: {ms:.2f} ms")
            if best is None or ms < best[0]:
                best = (ms, cand)
    return best[1]


def ray_query_leg(ctx, reps: int = 3, retune: bool = False) -> dict:
    """Ray-query microbenchmark: 1M random triangles in the unit cube (centroids uniform, sizes
    log-uniform in [1e-3, 1e-1]), 16M incoherent rays (origins uniform in the cube, directions
    uniform on the sphere) and 4096^2 coherent pinhole rays; closest hit id, t and barycentrics.
    Rays and hits are resident in HBM; the time is the closest-hit launch alone (HIP events on
    the launching stream), with each ray set's traversal settings (RAY_QUERY_TUNING)."""
    import numpy as np
    import torch
    from merian_amd import scenes
    pos, idx = scenes.ray_query(seed=1, n_tri=1_000_000)
    ds = ctx.triangles(pos, idx)
    n = 1 << 24
    d_inc = torch.from_numpy(scenes.incoherent_rays(n, seed=11).view(np.uint8)).cuda()
    d_coh = torch.empty(n * 32, dtype=torch.uint8, device="cuda")
    ctx.camera_rays_dev(ctx.camera(scenes.ray_query_camera(4096), 0), 4096, 4096, d_coh.data_ptr())
    d_hits = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    ctx.synchronize()
    stream = torch.cuda.Stream()
    res = {"workload": "1,000,000 random triangles in the unit cube; 16,777,216 incoherent rays "
                       "and 4096x4096 coherent pinhole rays"}
    for name, d_r in (("incoherent", d_inc), ("coherent", d_coh)):
        settings = _search_ray_tuning(ctx, ds, d_r, n, d_hits, stream) if retune else RAY_QUERY_TUNING[name]
        _apply(ds, settings)
        nodes, tris = ctx.trace_count_dev(ds, d_r.data_ptr(), n, d_hits.data_ptr())
        ms = _time_rays(ctx, ds, d_r, n, d_hits, stream, reps)
        bytes_per_ray = (nodes * NODE_BYTES[settings["bvh_width"]] + tris * TRI_BYTES) / n + RAY_IO_BYTES
        res[name] = {"mrays_per_s": round(n / ms / 1e3, 1), "ms_per_launch": round(ms, 3),
                     "node_visits_per_ray": round(nodes / n, 2), "triangle_tests_per_ray": round(tris / n, 2),
                     "algorithmic_gbs": round(bytes_per_ray * n / ms / 1e6, 1), "tuning": settings}
    ds.destroy()
    log(f"ray-query leg: {res}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--max-path-length", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--tune-ms", type=float, default=60000.0, help="autotuning budget of --retune")
    ap.add_argument("--retune", action="store_true",
                    help="search the tuning again instead of applying the cached autotune result")
    ap.add_argument("--default-tuning", action="store_true", help="render with the untuned defaults")
    ap.add_argument("--no-guided", action="store_true", help="skip the equal-time guided vs plain leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-mix leg")
    ap.add_argument("--no-ray-query", action="store_true", help="skip the ray-query microbenchmark leg")
    ap.add_argument("--no-small", action="store_true", help="skip the small plain-pt leg (512^2, 4 spp, L = 4)")
    ap.add_argument("--retune-rays", action="store_true",
                    help="grid-search the ray-query leg's traversal settings instead of using RAY_QUERY_TUNING")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from merian_amd import abi, scenes
    from merian_amd.runtime import Context, pt_params

    t0 = time.perf_counter()
    scene = scenes.full_scene(seed=2)
    spec = scene.camera
    W, H = spec.width, spec.height
    log(f"scene: {scene.triangle_count} triangles, {len(scene.materials)} materials ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.init()
    dev = torch.device("cuda:0")
    ctx = Context(0)
    t0 = time.perf_counter()
    ds = ctx.scene(scene)
    info = ds.info()
    log(f"BVH + upload: {time.perf_counter() - t0:.1f} s, {info}")
    gb = torch.empty(W * H * 16, dtype=torch.uint8, device=dev)
    out = torch.empty(W * H * 4, dtype=torch.float32, device=dev)
    params = pt_params(spp=args.spp, max_path_length=args.max_path_length, record_signature=False)
    tuning, tuning_source = None, "defaults"
    if args.retune:
        t0 = time.perf_counter()
        tuning = ds.tune(ctx.camera(spec, 0), W, H, params, budget_ms=args.tune_ms).as_dict()
        tuning_source = f"mrt_tune now ({time.perf_counter() - t0:.1f} s)"
        log(f"autotuned in {time.perf_counter() - t0:.1f} s: {tuning}")
    elif not args.default_tuning and os.path.exists(TUNING_CACHE) and args.spp == 64 and args.max_path_length == 8:
        with open(TUNING_CACHE) as f:
            cached = json.load(f)["tuning"]
        t = ds.tuning()
        for k, v in cached.items():
            setattr(t, k, v)
        ds.set_tuning(t)
        tuning = ds.tuning().as_dict()
        tuning_source = "cached mrt_tune result (merian_amd/tuning/full_1080p.json)"
        log(f"applied cached tuning {tuning}")

    def frame(f: int, count: bool = False):
        cam = ctx.camera(spec, f)
        ctx.gbuffer_dev(ds, cam, W, H, gb.data_ptr())
        params.count_traversal = int(count)
        return ctx.pt_render_dev(ds, params, cam, f, W, H, gb.data_ptr(), out.data_ptr(), stats=True)

    for k in range(args.warmup):
        frame(k)
        log(f"warmup {k} done")
    torch.cuda.synchronize()
    segs = 0
    ms_trace = 0.0
    launches = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = frame(args.warmup + k)
        segs += st.primary_segments + st.continuation_segments
        ms_trace += st.ms_trace
        launches += st.trace_launches
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = segs / dt / 1e6
    log(f"timed {args.steps} frames: {dt:.3f} s, {segs} segments, {value:.1f} Mseg/s")

    # PCIe-inclusive rate (reported beside `value`, never as it): the host-buffer entry point
    # mrt_render_frame, whose frame ends with the 33 MB irradiance copy into pageable host memory
    h_segs = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        cam_k = ctx.camera(spec, args.warmup + k)
        _, _, _, st_h = ctx.render(ds, params, cam_k, args.warmup + k, W, H, signature=False, gbuffer=False)
        h_segs += st_h.primary_segments + st_h.continuation_segments
    dt_h = time.perf_counter() - t0
    pcie = {"mseg_per_s": round(h_segs / dt_h / 1e6, 2), "ms_per_frame": round(dt_h * 1e3 / args.steps, 2),
            "entry_point": "mrt_render_frame (gbuffer + pt + irradiance copy to pageable host memory)"}
    log(f"host-buffer frames: {pcie}")

    # roofline of the dominant kernel (the intersect kernel): algorithmic bytes from the counting
    # variant on the first timed frame, scaled to the timed frames' continuation segments
    st_c = frame(args.warmup, count=True)
    cont = max(1, st_c.continuation_segments)
    node_bytes = NODE_BYTES[int(ds.tuning().bvh_width)]
    bytes_per_seg = (st_c.node_visits * node_bytes + st_c.triangle_tests * TRI_BYTES) / cont + RAY_IO_BYTES
    cont_timed = segs - args.steps * W * H
    bytes_per_launch = bytes_per_seg * cont_timed / max(1, launches)
    avg_ms = ms_trace / max(1, launches)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic, prof = None, {}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tj = json.load(f)
        # the profile holds for the tuning it was taken with
        prof_t = tj.get("tuning")
        same = (prof_t is None) == (tuning is None) and (
            prof_t is None or all(k in tuning and abs(float(prof_t[k]) - float(tuning[k])) < 1e-6 for k in prof_t))
        if same:
            traffic = tj["traffic_bytes_per_dispatch"]
            # measured HBM rate of the same kernel, and its issue figures, from the same profile
            prof = {"hbm_gbs": round(traffic / (avg_ms * 1e-3) / 1e9, 1),
                    "hbm_frac": round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "l2_hit_pct": tj.get("l2_hit_pct"), "valu_busy_pct": tj.get("valu_busy_pct"),
                    "valu_lane_util_pct": tj.get("valu_lane_util_pct"),
                    "profile": os.path.relpath(TRAFFIC_FILE, ROOT)}
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": round(bytes_per_launch),
            "kernel": "k_trace_queue", "avg_launch_ms": round(avg_ms, 4), "launches": launches,
            "bytes_per_segment": round(bytes_per_seg, 1), "node_bytes": node_bytes,
            "node_visits_per_segment": round(st_c.node_visits / cont, 2),
            "triangle_tests_per_segment": round(st_c.triangle_tests / cont, 2), **prof}

    line = {"metric": METRIC, "value": round(value, 2), "unit": "Mpath-segments/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / args.steps, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"full_1080p: {W}x{H}, {args.spp} spp, max path length {args.max_path_length}, "
                                   f"{scene.triangle_count} triangles (256 displaced tori + open room), "
                                   f"{len(scene.materials)} materials, 2048x1024 lat-long sky, BSDF-only pt",
                       "segments_per_step": segs // args.steps, "trace_ms_per_step": round(ms_trace / args.steps, 2),
                       "tuning": tuning, "tuning_source": tuning_source},
            "roofline": roof, "pcie_inclusive": pcie}
    ds.destroy()
    if not args.no_guided:
        line["guided"] = guided_leg(ctx, scene, args.max_path_length)
    if not args.no_serving:
        line["serving"] = serving_leg(ctx)
    if not args.no_ray_query:
        line["ray_query"] = ray_query_leg(ctx, retune=args.retune_rays)
    if not args.no_small:
        line["small"] = small_leg(ctx)
    ctx.close()
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads ...")
        line["cpu_baseline"] = cpu_baseline(scene, spec, args.warmup, args.cpu_seconds, threads)
        line["gpu_over_cpu"] = round(value / line["cpu_baseline"]["value"], 1)
        if "small" in line:
            cs = cpu_small(threads)
            line["cpu_baseline"]["small"] = cs
            line["small"]["gpu_over_cpu"] = round(line["small"]["mseg_per_s"] / cs["value"], 1)
        if "ray_query" in line:
            rq = cpu_ray_query(threads)
            line["cpu_baseline"]["ray_query"] = rq
            for k in ("incoherent", "coherent"):
                line["ray_query"][k]["gpu_over_cpu"] = round(line["ray_query"][k]["mrays_per_s"] / rq[k], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
