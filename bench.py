#!/usr/bin/env python
"""Benchmark: path segments per second of one pt frame on one MI355X.

Workload (the headline metric's configuration): 1920x1080, 64 spp, max path length 8, the seeded
procedural scene of 1,999,896 triangles (256 displaced tori instances in an enclosing room), 256
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
# passes over this workload, calibrated against gather and copy yardsticks (python -m merian_amd.profiling traffic)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_trace_queue_full_1080p.traffic.json")
TRAFFIC_WORKLOAD = "full_1080p enclosed room: 1920x1080, 64 spp, max path length 8, 1999896 triangles, BSDF-only pt"
# autotune result of the workload, applied before timing (re-run the search with --retune)
TUNING_CACHE = os.path.join(ROOT, "merian_amd", "tuning", "full_1080p.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def trace_variant(tuning: dict) -> int:
    """Template argument W of the intersect kernel the settings launch (4: BVH4; 8: BVH8 with
    sorted entries; 9: BVH8 in octant order with grouped stack entries)."""
    w = int(tuning.get("bvh_width", 4))
    return 9 if w == 8 and int(tuning.get("bvh8_order", 0)) == 1 else w


def roofline(kernel: str, avg_ms: float, launches: int, algorithmic_bytes: float, profile: str,
             match: dict | None = None, extra: dict | None = None) -> dict:
    """Bandwidth roofline of one kernel. `achieved` is the HBM-side traffic per dispatch that a
    rocprofv3 profile of the same workload measured (FETCH_SIZE + WRITE_SIZE, each scaled by what
    the counter reads for a yardstick of known size; profiles/*.traffic.json) over the kernel's
    average launch time measured here with HIP events: the bytes the kernel actually moves. The
    algorithmic bytes (every node, triangle and ray the traversal touches, most of them served by
    L2) stay beside it as a diagnostic. The profile applies only if it was taken with the settings
    in `match` (tuning, workload); otherwise `achieved` falls back to the algorithmic rate and says
    so; a profile of another kernel does not apply either. Reads are scaled by what FETCH_SIZE shows for random whole-line gathers of known size
    (exact: 64 bytes per request, two per 128-byte line fill); `traffic_upper` scales them by the
    copy's factor instead (streaming reads count one request per 128-byte line)."""
    alg_gbs = algorithmic_bytes / (avg_ms * 1e-3) / 1e9
    r = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "kernel": kernel, "avg_launch_ms": round(avg_ms, 4),
         "launches": launches, "algorithmic_bytes_per_launch": round(algorithmic_bytes),
         "algorithmic_gbs": round(alg_gbs, 1)}
    tj = None
    if os.path.exists(profile):
        with open(profile) as f:
            tj = json.load(f)
        for k, v in (match or {}).items():
            pv = tj.get(k)
            if isinstance(v, dict):
                if not isinstance(pv, dict) or any(kk not in v or abs(float(pv[kk]) - float(v[kk])) > 1e-6 for kk in pv):
                    tj = None
                    break
            elif pv != v:
                tj = None
                break
    # a profile of another kernel variant (e.g. the other BVH8 visiting order) does not apply
    if tj is not None and tj.get("kernel") not in (None, kernel):
        tj = None
    if tj is not None:
        traffic = float(tj["traffic_bytes_per_dispatch"])
        achieved = traffic / (avg_ms * 1e-3) / 1e9
        r.update({"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": round(traffic),
                  "basis": "HBM-side bytes per dispatch / launch time: FETCH_SIZE scaled by a random line-gather "
                           "yardstick, WRITE_SIZE by a copy" if "gather_factor_fetch" in tj else
                           "HBM-side bytes per dispatch (FETCH_SIZE + WRITE_SIZE, copy-calibrated) / launch time",
                  "traffic_over_algorithmic": round(traffic / algorithmic_bytes, 3),
                  "profile": os.path.relpath(profile, ROOT)})
        if tj.get("traffic_upper_bytes_per_dispatch") is not None:
            up = float(tj["traffic_upper_bytes_per_dispatch"])
            r.update({"traffic_upper": round(up), "frac_upper": round(up / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
        for k in ("l2_hit_pct", "valu_busy_pct", "valu_lane_util_pct", "mall_hit_pct", "dram_read_gbs"):
            if tj.get(k) is not None:
                r[k] = tj[k]
    else:
        r.update({"achieved": round(alg_gbs, 1), "frac": round(alg_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                  "basis": "algorithmic bytes (no profile of this workload and tuning)"})
    if extra:
        r.update(extra)
    return r


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


SMALL_TUNING = os.path.join(ROOT, "merian_amd", "tuning", "small_512.json")
SMALL_TRAFFIC = os.path.join(ROOT, "profiles", "k_trace_queue_small_512.traffic.json")
SMALL_WORKLOAD = "small: 8x8 icospheres (81,922 triangles), 512x512, 4 spp, max path length 4, constant environment"


def small_leg(ctx, frames: int = 64, retune: bool = False) -> dict:
    """The small plain-pt configuration: 8x8 icospheres on a ground plane (81,922 triangles),
    constant white environment, 512x512, 4 spp, max path length 4; a step = GBuffer + pt frame
    with its statistics read back. Settings: the cached mrt_tune result of this view
    (merian_amd/tuning/small_512.json; --retune searches again). Roofline of its intersect kernel
    from its own profile (profiles/k_trace_queue_small_512.traffic.json)."""
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
    if retune or not os.path.exists(SMALL_TUNING):
        ds.tune(ctx.camera(spec, 0), W, H, params, budget_ms=20000.0)  # mrt_tune on this view, untimed
        source = f"mrt_tune now ({time.perf_counter() - t0:.1f} s)"
    else:
        with open(SMALL_TUNING) as f:
            _apply(ds, json.load(f)["tuning"])
        source = "cached mrt_tune result (merian_amd/tuning/small_512.json)"
    tuning = {k: v for k, v in ds.tuning().as_dict().items() if k != "ms_per_frame"}

    def frame(f, count=False):
        cam = ctx.camera(spec, f)
        ctx.gbuffer_dev(ds, cam, W, H, gb.data_ptr())
        params.count_traversal = int(count)
        return ctx.pt_render_dev(ds, params, cam, f, W, H, gb.data_ptr(), out.data_ptr(), stats=True)

    frame(0)
    torch.cuda.synchronize()
    segs, ms_trace, launches = 0, 0.0, 0
    t0 = time.perf_counter()
    for f in range(1, frames + 1):
        st = frame(f)
        segs += st.primary_segments + st.continuation_segments
        ms_trace += st.ms_trace
        launches += st.trace_launches
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st_c = frame(1, count=True)
    cont = max(1, st_c.continuation_segments)
    node_bytes = NODE_BYTES[int(tuning["bvh_width"])]
    bytes_per_seg = (st_c.node_visits * node_bytes + st_c.triangle_tests * TRI_BYTES) / cont + RAY_IO_BYTES
    cont_timed = segs - frames * W * H
    avg_ms = ms_trace / max(1, launches)
    roof = roofline(f"k_trace<false, SrcQueue, SinkQueue, false, {trace_variant(tuning)}>", avg_ms, launches,
                    bytes_per_seg * cont_timed / max(1, launches), SMALL_TRAFFIC,
                    match={"tuning": tuning, "workload": SMALL_WORKLOAD},
                    extra={"bytes_per_segment": round(bytes_per_seg, 1)})
    ds.destroy()
    res = {"workload": f"{sc.triangle_count} triangles, 512x512, 4 spp, max path length 4, constant environment",
           "mseg_per_s": round(segs / dt / 1e6, 1), "ms_per_frame": round(dt * 1e3 / frames, 3),
           "segments_per_frame": segs // frames, "frames": frames, "tuning": tuning, "tuning_source": source,
           "roofline": roof}
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


def guided_leg(ctx, scene, L: int, spp: int = 64, ref_frames: int = 256, budget_s: float = 3.0,
               train_frames: int | None = None, log2_table: int = 14, guided_bounces: int = 4) -> dict:
    """Learned guiding at the 1080p workload's settings (1920x1080, 64 spp frames, L = 8): relMSE at
    equal device time against a 16384-spp plain render (256 frames of 64 spp).

    The reference's own noise is measured and subtracted: per pixel and channel its variance is the
    scatter of its frame means over the number of frames, and for an image x independent of the
    reference E[(x - ref)^2] = E[(x - mu)^2] + var(ref), so relMSE = mean(((x - ref)^2 - var(ref)) /
    (ref^2 + 1e-2)); the raw figures are reported beside it.

    Protocol: the guide trains online after each of its first `train_frames` frames (None: every
    frame; 16 Adam steps of the frame's records each) and then renders with the weights frozen;
    every frame's samples are kept (one-sample MIS keeps each
    frame unbiased whatever the weights). Frames render on one stream with device-side accumulation;
    device time is taken with HIP events around each frame. The plain pt and the Markov-chain
    guiding comparator (MCPG, learning online from an empty state with the MCPG example graph's
    grid settings) render frames until they have used the same device time."""
    import numpy as np
    import torch
    from merian_amd import abi, scenes
    from merian_amd.runtime import Guide, Mcpg, guide_config, mcpg_config, pt_params
    spec = scene.camera
    w, h = spec.width, spec.height
    ds = ctx.scene(scene)
    with open(TUNING_CACHE) as f:
        _apply(ds, json.load(f)["tuning"])
    cam = ctx.camera(spec, 0)
    strm = torch.cuda.current_stream().cuda_stream
    gb = torch.empty(w * h * 16, dtype=torch.uint8, device="cuda")
    out = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
    ctx.gbuffer_dev(ds, cam, w, h, gb.data_ptr())
    ctx.synchronize()

    def frame(f, params, guide=None, train=False, mcpg=None):
        if mcpg is not None:
            mcpg.render_dev(ds, params, cam, f, w, h, gb.data_ptr(), out.data_ptr(), stream=strm)
        elif guide is None:
            ctx.pt_render_dev(ds, params, cam, f, w, h, gb.data_ptr(), out.data_ptr(), stream=strm)
        else:
            guide.render_dev(ds, params, cam, f, w, h, gb.data_ptr(), out.data_ptr(), train=train, stream=strm)

    def run(params, base, budget=None, frames=None, **kw):
        """Frames until `frames` are done or `budget` seconds of device time are used; returns the
        mean image, frames and device seconds."""
        acc = torch.zeros(w * h * 4, dtype=torch.float64, device="cuda")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        used, n = 0.0, 0
        while (frames is None or n < frames) and (budget is None or used < budget):
            e0.record()
            frame(base + n, params, **kw)
            acc += out
            e1.record()
            e1.synchronize()
            used += e0.elapsed_time(e1) * 1e-3
            n += 1
        return (acc / n).cpu().numpy().reshape(h, w, 4), n, used

    plain = pt_params(spp=spp, max_path_length=L, record_signature=False)
    # reference: mean and per-pixel variance of the mean from the frames' scatter
    t0 = time.perf_counter()
    s1 = torch.zeros(w * h * 4, dtype=torch.float64, device="cuda")
    s2 = torch.zeros_like(s1)
    for f in range(ref_frames):
        frame(100000 + f, plain)
        o = out.double()
        s1 += o
        s2 += o * o
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    mean = s1 / ref_frames
    var_mean = (s2 / ref_frames - mean * mean).clamp_min(0) / (ref_frames - 1)
    ref = mean.cpu().numpy().reshape(h, w, 4)[..., :3]
    var_ref = var_mean.cpu().numpy().reshape(h, w, 4)[..., :3]
    del s1, s2, mean, var_mean
    den = ref ** 2 + 1e-2

    def rel(x):
        d = (x[..., :3] - ref) ** 2
        return float(np.mean(d / den)), float(np.mean((d - var_ref) / den))

    # guided: online training on the first frames, then frozen; every frame kept
    g = Guide(ctx, ds, guide_config(seed=1, train_steps=16, guided_bounces=guided_bounces, log2_table=log2_table))
    gp = pt_params(spp=spp, max_path_length=L, mode=abi.MODE_GUIDED, record_signature=False)
    gacc = torch.zeros(w * h * 4, dtype=torch.float64, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_g, n_g, t_train = 0.0, 0, 0.0
    while (train_frames is not None and n_g < train_frames) or t_g < budget_s:
        train = train_frames is None or n_g < train_frames
        e0.record()
        frame(1000 + n_g, gp, guide=g, train=train)
        gacc += out
        e1.record()
        e1.synchronize()
        dt = e0.elapsed_time(e1) * 1e-3
        t_g += dt
        t_train += dt if train else 0.0
        n_g += 1
    gimg = (gacc / n_g).cpu().numpy().reshape(h, w, 4)
    g.close()
    del gacc
    pimg, n_p, t_p = run(plain, 200000, budget=t_g)
    # the MCPG example graph's grid settings, counting guided directions only inside the BSDF
    # sampler's support (as guiding and NEE here do): with the reference's rule a guided
    # refraction where the GGX BTDF's pdf is 0 carries an unbounded weight
    m = Mcpg(ctx, mcpg_config(mc_tan_alpha_half=0.003, mc_distribution_dimension=1.5, lc_tan_alpha_half=0.012,
                              lc_buffer_size=8000009, lc_probe_count=2, bsdf_support_only=1))
    mimg, n_m, t_m = run(plain, 5000, budget=t_g, mcpg=m)
    m.close()
    ds.destroy()
    (rg_raw, rg), (rp_raw, rp), (rm_raw, rm) = rel(gimg), rel(pimg), rel(mimg)
    ref_noise = float(np.mean(var_ref / den))
    res = {"workload": f"{w}x{h}, {spp} spp per frame, L={L}, {scene.triangle_count:,} triangles (the 1080p scene); "
                       f"reference {ref_frames * spp} spp plain",
           "protocol": "equal device time (HIP events per frame): guided trains online after "
                       + ("every frame" if train_frames is None else f"its first {train_frames} frames, then frozen")
                       + f" (16 Adam steps each, guide on vertices 0-{guided_bounces - 1}, 2^{log2_table}-entry "
                       "hash tables), all frames kept; "
                       "plain and MCPG render frames until the same device time",
           "relmse_guided": rg, "relmse_plain": rp, "guided_over_plain": round(rg / rp, 4) if rp > 0 else None,
           "relmse_guided_raw": rg_raw, "relmse_plain_raw": rp_raw, "reference_noise_relmse": ref_noise,
           "seconds_guided": round(t_g, 3), "seconds_guided_training": round(t_train, 3),
           "seconds_plain": round(t_p, 3), "frames_guided": n_g, "frames_plain": n_p,
           "ms_per_frame_guided": round(1e3 * t_g / max(1, n_g), 2),
           "ms_per_frame_plain": round(1e3 * t_p / n_p, 2),
           "relmse_mcpg": rm, "relmse_mcpg_raw": rm_raw, "seconds_mcpg": round(t_m, 3), "frames_mcpg": n_m,
           "mcpg_note": "MCPG with bsdf_support_only (a guided direction counts only where the BSDF's sampling pdf "
                        "is > 0): with the reference's rule the GGX BTDF's non-zero evaluation outside that support "
                        "gives guided refractions unbounded weights",
           "reference_seconds": round(t_ref, 2)}
    log(f"guided leg: {res}")
    return res


MAX_PATHS_PER_JOB = 1 << 26  # paths in flight per serving job: a 1080p 64-spp job runs in 2 passes


def serving_leg(ctx, single_frame_mseg: float, n_jobs: int = 400, tune_ms: float = 1500.0, slots: int = 8) -> dict:
    """Serving mix: Poisson arrivals over 8 resident scenes (seeds 101-108, 103,682 to 1,999,896
    triangles), resolutions {256^2, 512^2, 1024x768, 1920x1080}, spp {1, 4, 16, 64}, L in 4..8,
    admitted by the server's HBM budget, at 80 % of the capacity a closed batch of 96 jobs of the
    mix reaches (submitted at once, timed until the last is done).
    Latency = submit -> result in the caller's host buffer; p50 / p99 are nearest-rank over the
    `n_jobs` timed jobs. A job's host buffer is dropped as soon as it is done."""
    import numpy as np
    from merian_amd import abi, scenes
    from merian_amd.runtime import Server, pt_params
    rng = np.random.default_rng(3)
    t0 = time.perf_counter()
    resident = scenes.serving_scenes(8)
    dss = [ctx.scene(s) for s in resident]
    log(f"serving: 8 scenes resident ({time.perf_counter() - t0:.1f} s): {[s.triangle_count for s in resident]}")
    res_choices = [(256, 256), (512, 512), (1024, 768), (1920, 1080)]
    t0 = time.perf_counter()
    for sc, d in zip(resident, dss):  # each scene autotuned once on a 1024x768, 16 spp, L = 6 view (untimed)
        spec = scenes.CameraSpec(sc.camera.eye, sc.camera.target, sc.camera.up, sc.camera.vfov, 1024, 768)
        d.tune(ctx.camera(spec, 0), 1024, 768, pt_params(spp=16, max_path_length=6, record_signature=False),
               budget_ms=tune_ms)
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

    with Server(ctx, max_slots=slots, max_paths_per_job=MAX_PATHS_PER_JOB) as srv:
        # capacity: a closed batch of 96 jobs of the mix, submitted at once, over the time until
        # the last one is done (the drain of the last jobs included, so the Poisson arrivals at
        # 80 % of it keep the queue stable)
        batch = [job() for _ in range(96)]
        t0 = time.perf_counter()
        for h_ in [srv.submit(*j) for j in batch]:
            h_.wait()
        capacity = len(batch) / (time.perf_counter() - t0)
        cap_st = srv.stats()
        cap_mseg = cap_st.segments / (time.perf_counter() - t0) / 1e6
        srv.reset_stats()
        lam = 0.8 * capacity
        jobs = [job() for _ in range(n_jobs)]
        arrivals = np.cumsum(rng.exponential(1.0 / lam, len(jobs)))
        pending = []
        t0 = time.perf_counter()
        t_sweep = 0.0
        for a, j in zip(arrivals, jobs):
            now = time.perf_counter() - t0
            if now - t_sweep > 0.05:
                # drop finished jobs (and their host buffers) every 50 ms: polling every pending
                # job in a tight loop would contend for the server's lock with its scheduler
                pending = [p_ for p_ in pending if p_.poll().state < 2]
                t_sweep = now
            dt = a - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            pending.append(srv.submit(*j))
        for p_ in pending:
            p_.wait()
        wall = time.perf_counter() - t0
        st = srv.stats()
    for d in dss:
        d.destroy()
    mseg = st.segments / wall / 1e6
    res = {"workload": f"{n_jobs} Poisson-arrival jobs over 8 resident scenes "
                       f"({', '.join(f'{s.triangle_count:,}' for s in resident)} triangles), res {{256^2, 512^2, "
                       f"1024x768, 1920x1080}}, spp {{1,4,16,64}}, L 4..8, {slots} slots, "
                       f"{MAX_PATHS_PER_JOB} paths in flight per job",
           "arrival_rate_per_s": round(lam, 2), "capacity_jobs_per_s": round(capacity, 2),
           "capacity_mseg_per_s": round(cap_mseg, 1),
           "capacity_fraction_of_single_frame": round(cap_mseg / single_frame_mseg, 3),
           "p50_ms": round(st.p50_ms, 2), "p99_ms": round(st.p99_ms, 2), "max_ms": round(st.max_ms, 2),
           "mean_ms": round(st.mean_ms, 2), "jobs_per_s": round(st.completed / wall, 2),
           "mseg_per_s": round(mseg, 1), "fraction_of_single_frame": round(mseg / single_frame_mseg, 3),
           "peak_in_flight": st.peak_in_flight, "completed": st.completed, "wall_s": round(wall, 2),
           "tuning_source": f"mrt_tune per scene before the mix ({tune_s:.1f} s)"}
    log(f"serving leg: {res}")
    return res


# Traversal settings of the ray-query leg, per ray set: the fastest of a grid search over BVH
# width {4, 8}, leaf size {1, 2, 4}, SAH node cost {0.5, 1, 2} (BVH4 leaf 2 only), refill lanes
# {8, 16, 24, 32, 48}, leaf-phase lanes {4, 8, 16, 32} and grid {50, 75, 100} % on this workload
# (re-run with --retune-rays). No setting changes a hit.
RAY_QUERY_TUNING = {
    "incoherent": dict(bvh_width=8, bvh_max_leaf=2, bvh_traversal_cost=1.0, refill_lanes=16, leaf_lanes=4,
                       grid_percent=100, bvh8_order=1),
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
    # (width, leaf, SAH cost, BVH8 visiting order)
    trees = [(4, 2, 1.0, 0), (4, 1, 1.0, 0), (4, 4, 1.0, 0), (8, 2, 1.0, 0), (8, 4, 1.0, 0), (8, 1, 1.0, 1),
             (8, 2, 1.0, 1), (8, 4, 1.0, 1), (4, 2, 2.0, 0), (4, 2, 0.5, 0)]
    for width, leaf, cost, order in trees:
        _apply(ds, dict(bvh_width=width, bvh_max_leaf=leaf, bvh_traversal_cost=cost, bvh8_order=order))
        for refill, ll, grid in itertools.product([8, 16, 24, 32, 48], [4, 8, 16, 32], [50, 75, 100]):
            cand = dict(bvh_width=width, bvh_max_leaf=leaf, bvh_traversal_cost=cost, refill_lanes=refill,
                        leaf_lanes=ll, grid_percent=grid, bvh8_order=order)
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
        roof = roofline(f"k_trace<false, SrcRays, SinkHits, false, {trace_variant(settings)}>", ms, reps,
                        bytes_per_ray * n, os.path.join(ROOT, "profiles", f"k_trace_rays_{name}.traffic.json"),
                        match={"tuning": settings, "workload": f"ray_query {name}"})
        res[name] = {"mrays_per_s": round(n / ms / 1e3, 1), "ms_per_launch": round(ms, 3),
                     "node_visits_per_ray": round(nodes / n, 2), "triangle_tests_per_ray": round(tris / n, 2),
                     "algorithmic_gbs": round(bytes_per_ray * n / ms / 1e6, 1), "tuning": settings,
                     "roofline": roof}
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
    ap.add_argument("--no-nee", action="store_true", help="skip the NEE frames of the 1080p workload")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-mix leg")
    ap.add_argument("--serving-jobs", type=int, default=400, help="timed jobs of the serving mix")
    ap.add_argument("--serving-slots", type=int, default=8, help="jobs in flight in the serving mix")
    ap.add_argument("--single-frame-mseg", type=float, default=0.0,
                    help="--only serving: the single-frame rate its fractions are quoted against")
    ap.add_argument("--no-ray-query", action="store_true", help="skip the ray-query microbenchmark leg")
    ap.add_argument("--no-small", action="store_true", help="skip the small plain-pt leg (512^2, 4 spp, L = 4)")
    ap.add_argument("--retune-rays", action="store_true",
                    help="grid-search the ray-query leg's traversal settings instead of using RAY_QUERY_TUNING")
    ap.add_argument("--only", choices=["guided", "serving", "rays", "small"],
                    help="run one leg alone and print its JSON (profiling passes)")
    ap.add_argument("--small-frames", type=int, default=64)
    ap.add_argument("--ray-reps", type=int, default=3)
    ap.add_argument("--ref-frames", type=int, default=256, help="64-spp frames of the guided leg's reference")
    ap.add_argument("--guide-table", type=int, default=14, help="log2 entries per hash-grid level of the guide")
    ap.add_argument("--guided-bounces", type=int, default=4, help="vertices [0, n) sample the guide")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from merian_amd import abi, scenes
    from merian_amd.runtime import Context, pt_params

    if args.only:
        torch.cuda.init()
        ctx = Context(0)
        if args.only == "small":
            leg = small_leg(ctx, frames=args.small_frames, retune=args.retune)
        elif args.only == "rays":
            leg = ray_query_leg(ctx, reps=args.ray_reps, retune=args.retune_rays)
        elif args.only == "serving":
            leg = serving_leg(ctx, args.single_frame_mseg or 1.0, n_jobs=args.serving_jobs, slots=args.serving_slots)
        else:
            leg = guided_leg(ctx, scenes.full_scene(seed=2), args.max_path_length, ref_frames=args.ref_frames,
                             log2_table=args.guide_table, guided_bounces=args.guided_bounces)
        ctx.close()
        print(json.dumps({args.only: leg}), flush=True)
        return

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

    # NEE mode on the same frames (environment importance sampling with MIS, one shadow ray per
    # vertex; no counterpart in the reference's pt): reported beside `value`, the BSDF-only pt
    p_nee = pt_params(spp=args.spp, max_path_length=args.max_path_length, mode=abi.MODE_NEE, record_signature=False)
    n_segs, n_shadow = 0, 0
    nee = None
    frame_nee = lambda f: ctx.pt_render_dev(ds, p_nee, ctx.camera(spec, f), f, W, H, gb.data_ptr(), out.data_ptr(),
                                            stats=True)
    if not args.no_nee:
        ctx.gbuffer_dev(ds, ctx.camera(spec, args.warmup), W, H, gb.data_ptr())
        frame_nee(args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            ctx.gbuffer_dev(ds, ctx.camera(spec, args.warmup + k), W, H, gb.data_ptr())
            st_n = frame_nee(args.warmup + k)
            n_segs += st_n.primary_segments + st_n.continuation_segments
            n_shadow += st_n.shadow_rays
        torch.cuda.synchronize()
        dt_n = time.perf_counter() - t0
        nee = {"mseg_per_s": round(n_segs / dt_n / 1e6, 1), "ms_per_frame": round(dt_n * 1e3 / args.steps, 2),
               "shadow_rays_per_frame": n_shadow // args.steps,
               "note": "path segments per second as `value`; shadow rays are traced on top and not counted"}
        log(f"NEE frames: {nee}")

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
    roof = roofline(f"k_trace<false, SrcQueue, SinkQueue, false, {trace_variant(ds.tuning().as_dict())}>", avg_ms, launches,
                    bytes_per_launch, TRAFFIC_FILE,
                    match={"tuning": tuning, "workload": TRAFFIC_WORKLOAD},
                    extra={"bytes_per_segment": round(bytes_per_seg, 1), "node_bytes": node_bytes,
                           "node_visits_per_segment": round(st_c.node_visits / cont, 2),
                           "triangle_tests_per_segment": round(st_c.triangle_tests / cont, 2)})

    line = {"metric": METRIC, "value": round(value, 2), "unit": "Mpath-segments/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / args.steps, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"full_1080p: {W}x{H}, {args.spp} spp, max path length {args.max_path_length}, "
                                   f"{scene.triangle_count} triangles (256 displaced tori in an enclosing room with a window and a skylight), "
                                   f"{len(scene.materials)} materials, 2048x1024 lat-long sky, BSDF-only pt",
                       "segments_per_step": segs // args.steps, "trace_ms_per_step": round(ms_trace / args.steps, 2),
                       "tuning": tuning, "tuning_source": tuning_source},
            "roofline": roof, "pcie_inclusive": pcie}
    if nee is not None:
        line["nee"] = nee
    ds.destroy()
    if not args.no_guided:
        line["guided"] = guided_leg(ctx, scene, args.max_path_length, ref_frames=args.ref_frames,
                                    log2_table=args.guide_table, guided_bounces=args.guided_bounces)
    if not args.no_serving:
        line["serving"] = serving_leg(ctx, value, n_jobs=args.serving_jobs, slots=args.serving_slots)
    if not args.no_ray_query:
        line["ray_query"] = ray_query_leg(ctx, reps=args.ray_reps, retune=args.retune_rays)
    if not args.no_small:
        line["small"] = small_leg(ctx, frames=args.small_frames, retune=args.retune)
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
