#!/usr/bin/env python
"""Benchmark: path segments per second of one pt frame on one MI355X.

Workload (the headline metric's configuration): 1920x1080, 64 spp, max path length 8, the seeded
procedural scene of 1,999,880 triangles (256 displaced tori instances + an open room), 256
materials (Lambert / GGX conductor / dielectric 1.5), 2048x1024 lat-long sky with a sun of
5e4:1 dynamic range. A step = one frame: GBuffer pass + pt (BSDF-only estimator, the
reference's). Inputs (scene, BVH, buffers) are resident in HBM before timing starts.

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
NODE_BYTES = 128      # BVH4Node
TRI_BYTES = 48        # TriRecord
RAY_IO_BYTES = 32 + 16  # queued ray in, hit out
# HBM bytes per dispatch of the intersect kernel, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes over this workload and calibrated against a plain copy (python -m merian_amd.profiling traffic)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_trace_queue_full_1080p.traffic.json")


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--max-path-length", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
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

    # roofline of the dominant kernel (the intersect kernel): algorithmic bytes from the counting
    # variant on the first timed frame, scaled to the timed frames' continuation segments
    st_c = frame(args.warmup, count=True)
    cont = max(1, st_c.continuation_segments)
    bytes_per_seg = (st_c.node_visits * NODE_BYTES + st_c.triangle_tests * TRI_BYTES) / cont + RAY_IO_BYTES
    cont_timed = segs - args.steps * W * H
    bytes_per_launch = bytes_per_seg * cont_timed / max(1, launches)
    avg_ms = ms_trace / max(1, launches)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f)["traffic_bytes_per_dispatch"]
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": round(bytes_per_launch),
            "kernel": "k_trace_queue", "avg_launch_ms": round(avg_ms, 4), "launches": launches,
            "bytes_per_segment": round(bytes_per_seg, 1),
            "node_visits_per_segment": round(st_c.node_visits / cont, 2),
            "triangle_tests_per_segment": round(st_c.triangle_tests / cont, 2)}

    line = {"metric": METRIC, "value": round(value, 2), "unit": "Mpath-segments/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / args.steps, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"full_1080p: {W}x{H}, {args.spp} spp, max path length {args.max_path_length}, "
                                   f"{scene.triangle_count} triangles (256 displaced tori + open room), "
                                   f"{len(scene.materials)} materials, 2048x1024 lat-long sky, BSDF-only pt",
                       "segments_per_step": segs // args.steps, "trace_ms_per_step": round(ms_trace / args.steps, 2)},
            "roofline": roof}
    ds.destroy()
    ctx.close()
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads ...")
        line["cpu_baseline"] = cpu_baseline(scene, spec, args.warmup, args.cpu_seconds, threads)
        line["gpu_over_cpu"] = round(value / line["cpu_baseline"]["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
