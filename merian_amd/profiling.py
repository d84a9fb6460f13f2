"""rocprofv3 helpers: a copy workload of known size (the yardstick for the FETCH_SIZE and
WRITE_SIZE counters) and summaries of rocprofv3's CSV output.

    python -m merian_amd.profiling copy --mib 1024 --reps 4
        one plain device-to-device copy of `mib` MiB per rep (run under rocprofv3 --pmc)
    python -m merian_amd.profiling gather --mib 4096 --rows 16777216 --reps 2
        random 128-byte line gathers of known count (run under rocprofv3 --pmc)
    python -m merian_amd.profiling stats DIR
        per-kernel table from DIR's *_kernel_stats.csv (rocprofv3 --kernel-trace --stats)
    python -m merian_amd.profiling pmc DIR COUNTER
        per-kernel mean of COUNTER per dispatch from DIR's *_counter_collection.csv
    python -m merian_amd.profiling traffic FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR KERNEL
            [--gather-fetch-dir DIR]
        calibrated HBM bytes per dispatch of KERNEL (JSON)
    python -m merian_amd.profiling valu DIR
        per-kernel VALU busy, lane utilisation and issue time from a pass with GRBM_GUI_ACTIVE,
        SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU and SQ_THREAD_CYCLES_VALU
    python -m merian_amd.profiling l2 DIR
        per-kernel L2 hit rate from a TCC_HIT_sum / TCC_MISS_sum pass
    python -m merian_amd.profiling l1 DIR
        per-kernel L1-to-L2 read requests (count, rate, L1 hit share) and texture-addresser busy
        and stall shares, from a pass with TCP_TCC_READ_REQ_sum, TCP_TOTAL_CACHE_ACCESSES_sum,
        TA_BUSY_avr, TA_ADDR_STALLED_BY_TC_CYCLES_sum and GRBM_GUI_ACTIVE
    python -m merian_amd.profiling dram DIR
        per-kernel L2-to-fabric requests and the share destined for DRAM, from a pass with
        TCC_EA0_RDREQ_sum, TCC_EA0_RDREQ_DRAM_sum, TCC_EA0_WRREQ_sum and TCC_EA0_WRREQ_DRAM_sum

gfx950's rocprofv3 lists no MALL (Infinity Cache) counter: the L2's fabric requests are the last
level it can see, and a request "destined for DRAM" still passes the MALL on its way.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict


def _one(pattern: str) -> str:
    files = sorted(glob.glob(pattern, recursive=True))
    if not files:
        raise FileNotFoundError(pattern)
    return files[0]


def kernel_stats(d: str) -> list[dict]:
    """Rows of rocprofv3's kernel statistics (name, calls, total/average ns, share)."""
    rows = []
    with open(_one(os.path.join(d, "**", "*kernel_stats.csv"))) as f:
        for r in csv.DictReader(f):
            rows.append({"name": r["Name"], "calls": int(r["Calls"]), "total_ns": float(r["TotalDurationNs"]),
                         "avg_ns": float(r["AverageNs"]), "pct": float(r["Percentage"])})
    rows.sort(key=lambda r: -r["total_ns"])
    return rows


def counter_per_kernel(d: str, counter: str) -> dict[str, dict]:
    """Mean of `counter` per dispatch, and dispatch count, per kernel name."""
    acc = defaultdict(lambda: [0.0, 0])
    with open(_one(os.path.join(d, "**", "*counter_collection.csv"))) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            a = acc[r["Kernel_Name"]]
            a[0] += float(r["Counter_Value"])
            a[1] += 1
    return {k: {"mean": v[0] / v[1], "dispatches": v[1]} for k, v in acc.items()}


def short_name(name: str) -> str:
    """Kernel name without return type, namespaces and parameter list (template arguments kept)."""
    n = name.replace("void ", "").replace("(anonymous namespace)::", "").replace("mrt::", "")
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return n[:i]
    return n


def traffic(fetch_dir: str, write_dir: str, copy_fetch_dir: str, copy_write_dir: str, kernel: str,
            copy_bytes: int, gather_fetch_dir: str = "", gather_bytes: int = 0) -> dict:
    """HBM-side bytes per dispatch of the kernels whose short name starts with `kernel`, from a
    FETCH_SIZE run and a WRITE_SIZE run.

    Yardsticks: a plain copy of `copy_bytes` (reads and writes exactly that many) and, when
    given, a random gather of whole 128-byte lines (`gather_bytes` of lines). On gfx950
    FETCH_SIZE counts 64 bytes per L2 read request: exact for scattered line fills (two requests
    per 128-byte line; the gather reads back its line bytes), half the bytes of the copy's
    streaming reads (one request per line). Reads are scaled by the gather's factor when it is
    given (the traversal's reads are scattered node and triangle fetches; its ray-queue stream
    would count low, so `traffic_upper_bytes_per_dispatch` scales the reads by the copy's factor
    instead), else by the copy's. Writes are scaled by the copy's write factor."""
    def sel(d, counter):
        tot, n = 0.0, 0
        for k, v in counter_per_kernel(d, counter).items():
            if short_name(k).startswith(kernel):
                tot += v["mean"] * v["dispatches"]
                n += v["dispatches"]
        if n == 0:
            raise ValueError(f"no dispatch of {kernel} in {d}")
        return tot / n, n

    def copy_kib(d, counter):
        vals = [v["mean"] for k, v in counter_per_kernel(d, counter).items() if "copy" in k.lower()]
        return max(vals)

    def gather_kib(d):
        vals = [v["mean"] for k, v in counter_per_kernel(d, "FETCH_SIZE").items() if "gather" in k.lower()]
        return max(vals)

    f_kib, nf = sel(fetch_dir, "FETCH_SIZE")
    w_kib, nw = sel(write_dir, "WRITE_SIZE")
    cf = copy_bytes / (copy_kib(copy_fetch_dir, "FETCH_SIZE") * 1024.0)
    cw = copy_bytes / (copy_kib(copy_write_dir, "WRITE_SIZE") * 1024.0)
    r = {"kernel": kernel, "dispatches": [nf, nw], "fetch_kib_per_dispatch": round(f_kib, 1),
         "write_kib_per_dispatch": round(w_kib, 1), "copy_bytes": copy_bytes,
         "copy_factor_fetch": round(cf, 4), "copy_factor_write": round(cw, 4)}
    gf = cf
    if gather_fetch_dir:
        gf = gather_bytes / (gather_kib(gather_fetch_dir) * 1024.0)
        r.update({"gather_line_bytes": gather_bytes, "gather_factor_fetch": round(gf, 4)})
    r["traffic_bytes_per_dispatch"] = round(f_kib * 1024.0 * gf + w_kib * 1024.0 * cw)
    r["traffic_upper_bytes_per_dispatch"] = round(f_kib * 1024.0 * max(cf, gf) + w_kib * 1024.0 * cw)
    return r


def agent(d: str) -> dict:
    """The GPU agent's row of rocprofv3's agent info (CU, SIMD and XCC counts, engine clock)."""
    with open(_one(os.path.join(d, "**", "*agent_info.csv"))) as f:
        for r in csv.DictReader(f):
            if r.get("Agent_Type", "").upper().find("GPU") >= 0 or r.get("Name", "").startswith("gfx"):
                return r
    raise ValueError(f"no GPU agent in {d}")


def _sums(d: str) -> tuple[dict, dict]:
    """Per kernel: counter sums over its dispatches, and (dispatches, summed duration in ms)."""
    acc = defaultdict(lambda: defaultdict(float))
    disp = defaultdict(dict)
    with open(_one(os.path.join(d, "**", "*counter_collection.csv"))) as f:
        for r in csv.DictReader(f):
            k = short_name(r["Kernel_Name"])
            acc[k][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[k][r["Dispatch_Id"]] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
    return acc, {k: (len(v), sum(v.values())) for k, v in disp.items()}


def valu(d: str) -> dict:
    """VALU figures per kernel. busy = active VALU cycles per CU over the GPU's busy cycles
    (GRBM_GUI_ACTIVE summed over the XCDs, so divided by their number); lane utilisation =
    thread-cycles / (64 x active cycles); issue_ms = wave64 VALU instructions x 4 cycles over
    all SIMDs at the peak engine clock."""
    a = agent(d)
    cus, simds, xcc = int(a["Cu_Count"]), int(a["Simd_Count"]), int(a["Num_Xcc"])
    clk_hz = float(a["Max_Engine_Clk_Fcompute"]) * 1e6
    acc, disp = _sums(d)
    out = {}
    for k, c in acc.items():
        n, ms = disp[k]
        if c.get("SQ_ACTIVE_INST_VALU", 0) <= 0 or c.get("GRBM_GUI_ACTIVE", 0) <= 0:
            continue
        gui = c["GRBM_GUI_ACTIVE"] / xcc
        out[k] = {"dispatches": n, "avg_ms": round(ms / n, 4),
                  "valu_insts_per_dispatch": round(c.get("SQ_INSTS_VALU", 0) / n),
                  "valu_busy_pct": round(100 * c["SQ_ACTIVE_INST_VALU"] / cus / gui, 1),
                  "valu_lane_util_pct": round(100 * c.get("SQ_THREAD_CYCLES_VALU", 0) / (64 * c["SQ_ACTIVE_INST_VALU"]), 1),
                  "valu_issue_ms": round(c.get("SQ_INSTS_VALU", 0) / n * 4 / simds / clk_hz * 1e3, 4)}
    return out


def l1(d: str) -> dict:
    """L1 (TCP) figures per kernel: read requests it sends to the L2 per dispatch and per second
    of the kernel's own duration, the share of its accesses that did not become L2 reads, the
    texture addresser's busy share and the share of cycles it stalled on the L1 (both per CU,
    over the GPU's busy cycles)."""
    a = agent(d)
    cus, xcc = int(a["Cu_Count"]), int(a["Num_Xcc"])
    acc, disp = _sums(d)
    out = {}
    for k, c in acc.items():
        n, ms = disp[k]
        req = c.get("TCP_TCC_READ_REQ_sum", 0.0)
        gui = c.get("GRBM_GUI_ACTIVE", 0.0) / xcc
        if req <= 0 or gui <= 0:
            continue
        tot = c.get("TCP_TOTAL_CACHE_ACCESSES_sum", 0.0)
        out[k] = {"dispatches": n, "avg_ms": round(ms / n, 4), "l2_read_req_per_dispatch": round(req / n),
                  "l2_read_greq_per_s": round(req / (ms * 1e-3) / 1e9, 1),
                  "l1_hit_pct": round(100 * (1 - req / tot), 1) if tot > 0 else None,
                  "ta_busy_pct": round(100 * c.get("TA_BUSY_avr", 0.0) / gui, 1),
                  "ta_stalled_on_l1_pct": round(100 * c.get("TA_ADDR_STALLED_BY_TC_CYCLES_sum", 0.0) / cus / gui, 1)}
    return out


def mfma(d: str) -> dict:
    """MFMA figures per kernel from a pass with SQ_INSTS_VALU_MFMA_MOPS_F16 (F16 MFMA work in
    units of 512 flops), SQ_VALU_MFMA_BUSY_CYCLES and GRBM_GUI_ACTIVE: achieved fp16 TFLOP/s over
    the kernel's own duration, its fraction of the 2.5 PFLOP/s dense fp16 peak, and MFMA busy =
    busy cycles per SIMD over the GPU's busy cycles."""
    a = agent(d)
    simds, xcc = int(a["Simd_Count"]), int(a["Num_Xcc"])
    acc, disp = _sums(d)
    out = {}
    for k, c in acc.items():
        n, ms = disp[k]
        mops = c.get("SQ_INSTS_VALU_MFMA_MOPS_F16", 0.0)
        if mops <= 0:
            continue
        tflops = mops * 512.0 / (ms * 1e-3) / 1e12
        gui = c.get("GRBM_GUI_ACTIVE", 0.0) / xcc
        busy = c.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0)
        out[k] = {"dispatches": n, "avg_ms": round(ms / n, 4), "fp16_gflop_per_dispatch": round(mops * 512 / n / 1e9, 3),
                  "fp16_tflops": round(tflops, 1), "frac_of_2500": round(tflops / 2500.0, 4),
                  "mfma_busy_pct": round(100 * busy / simds / gui, 1) if gui > 0 else None}
    return out


def guided_workload(width: int, height: int, spp: int, train_frames: int, frames: int, guided_bounces: int) -> None:
    """Learned-guide frames on a view of the 1080p scene: online-training frames (forward,
    backward, weight gradients, Adam), then frames with the guide frozen (inference only)."""
    import torch

    from . import abi as _abi
    from . import scenes
    from .runtime import Context, Guide, guide_config, pt_params
    ctx = Context()
    scene = scenes.full_scene()
    spec = scenes.CameraSpec(scene.camera.eye, scene.camera.target, scene.camera.up, scene.camera.vfov, width, height)
    ds = ctx.scene(scene)
    # the 1080p workload's cached settings (as bench.py applies them)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "full_1080p.json")) as f:
        t = ds.tuning()
        for k, v in json.load(f)["tuning"].items():
            setattr(t, k, v)
        ds.set_tuning(t)
    cam = ctx.camera(spec, 0)
    gb = torch.empty(width * height * 16, dtype=torch.uint8, device="cuda")
    out = torch.empty(width * height * 4, dtype=torch.float32, device="cuda")
    ctx.gbuffer_dev(ds, cam, width, height, gb.data_ptr())
    g = Guide(ctx, ds, guide_config(seed=1, train_steps=16, guided_bounces=guided_bounces))
    p = pt_params(spp=spp, max_path_length=8, mode=_abi.MODE_GUIDED, record_signature=False)
    for f in range(train_frames + frames):
        train = f < train_frames
        st = g.render_dev(ds, p, cam, f, width, height, gb.data_ptr(), out.data_ptr(), train=train, stats=True)
        torch.cuda.synchronize()
        print(json.dumps({"frame": f, "train": train, "ms": round(st.ms_total, 3),
                          "segments": st.continuation_segments}), flush=True)
    st = ctx.pt_render_dev(ds, pt_params(spp=spp, max_path_length=8, record_signature=False), cam, 0, width, height,
                           gb.data_ptr(), out.data_ptr(), stats=True)
    torch.cuda.synchronize()
    print(json.dumps({"plain_ms": round(st.ms_total, 3), "segments": st.continuation_segments}), flush=True)
    g.close()
    ds.destroy()
    ctx.close()


def l2(d: str) -> dict:
    """L2 hit rate per kernel (TCC_HIT / (TCC_HIT + TCC_MISS))."""
    acc, disp = _sums(d)
    return {k: {"dispatches": disp[k][0], "l2_hit_pct": round(100 * c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"]), 1)}
            for k, c in acc.items() if c.get("TCC_HIT_sum", 0) + c.get("TCC_MISS_sum", 0) > 0}


def dram(d: str) -> dict:
    """L2-to-fabric (EA) read and write requests per dispatch and the percentage of them
    destined for DRAM (the memory controller, behind the MALL) rather than IO/GMI."""
    acc, disp = _sums(d)
    out = {}
    for k, c in acc.items():
        n, ms = disp[k]
        rd, wr = c.get("TCC_EA0_RDREQ_sum", 0.0), c.get("TCC_EA0_WRREQ_sum", 0.0)
        if rd + wr <= 0:
            continue
        out[k] = {"dispatches": n, "avg_ms": round(ms / n, 4), "ea_rdreq_per_dispatch": round(rd / n),
                  "ea_wrreq_per_dispatch": round(wr / n),
                  "rd_to_dram_pct": round(100 * c.get("TCC_EA0_RDREQ_DRAM_sum", 0.0) / rd, 1) if rd > 0 else None,
                  "wr_to_dram_pct": round(100 * c.get("TCC_EA0_WRREQ_DRAM_sum", 0.0) / wr, 1) if wr > 0 else None}
    return out


def traffic_record(args) -> dict:
    """The traffic of one kernel (traffic()) with the workload and tuning it was measured on and
    the same kernel's figures from the other passes (kernel-trace average, VALU, L2, DRAM
    requests), as bench.py's roofline reads it."""
    r = traffic(args.fetch_dir, args.write_dir, args.copy_fetch_dir, args.copy_write_dir, args.kernel,
                args.copy_mib << 20, args.gather_fetch_dir, args.gather_rows * 128)
    if args.workload:
        r["workload"] = args.workload
    if args.tuning_file:
        with open(args.tuning_file) as f:
            r["tuning"] = json.load(f)["tuning"]
    def pick(rows):
        hits = [v for k, v in rows.items() if k.startswith(args.kernel)]
        return hits[0] if len(hits) == 1 else None
    if args.stats_dir:
        ks = [x for x in kernel_stats(args.stats_dir) if short_name(x["name"]).startswith(args.kernel)]
        if len(ks) == 1:
            r["avg_ms_kernel_trace"] = round(ks[0]["avg_ns"] / 1e6, 4)
            r["hbm_gbs_at_kernel_trace_avg"] = round(r["traffic_bytes_per_dispatch"] / (ks[0]["avg_ns"] * 1e-9) / 1e9, 1)
    if args.valu_dir and (v := pick(valu(args.valu_dir))):
        r.update({k: v[k] for k in ("valu_busy_pct", "valu_lane_util_pct", "valu_issue_ms")})
    if args.l2_dir and (v := pick(l2(args.l2_dir))):
        r["l2_hit_pct"] = v["l2_hit_pct"]
    if args.dram_dir and (v := pick(dram(args.dram_dir))):
        r.update({k: v[k] for k in ("ea_rdreq_per_dispatch", "ea_wrreq_per_dispatch", "rd_to_dram_pct",
                                    "wr_to_dram_pct")})
    if args.command:
        r["commands"] = args.command
    return r


def copy_workload(mib: int, reps: int) -> None:
    import torch
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    src.fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "reps": reps}), flush=True)


def gather_workload(mib: int, rows: int, reps: int) -> None:
    """Random 128-byte row gathers from a buffer of `mib` MiB (far larger than L2 and MALL): each
    rep reads `rows` rows at random 128-byte-aligned offsets (out = table[idx]), one whole
    128-byte L2 line each. A yardstick for the counters under scattered reads, as the traversal
    makes them: the lines read are known (`rows` x 128 bytes, plus the `rows` x 8-byte index
    stream)."""
    import torch
    table = torch.empty((mib << 20) // 128, 32, dtype=torch.float32, device="cuda:0")
    table.fill_(1.0)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(1)
    idx = torch.randint(0, table.shape[0], (rows,), device="cuda:0", generator=g)
    torch.cuda.synchronize()
    for _ in range(reps):
        out = table.index_select(0, idx)
    torch.cuda.synchronize()
    print(json.dumps({"table_bytes": table.numel() * 4, "rows": rows, "row_bytes": 128, "reps": reps,
                      "checksum": float(out.sum())}), flush=True)


def mcpg_workload(width: int, height: int, spp: int, frames: int, max_path_length: int) -> None:
    """MCPG frames on a view of the 1080p scene (the MCPG example graph's grid settings), then
    one plain frame of the same view: a workload to profile the MCPG kernels."""
    import torch

    from . import scenes
    from .runtime import Context, Mcpg, mcpg_config, pt_params
    ctx = Context()
    scene = scenes.full_scene()
    spec = scenes.CameraSpec(scene.camera.eye, scene.camera.target, scene.camera.up, scene.camera.vfov, width, height)
    ds = ctx.scene(scene)
    cam = ctx.camera(spec, 0)
    gb = torch.empty(width * height * 16, dtype=torch.uint8, device="cuda")
    out = torch.empty(width * height * 4, dtype=torch.float32, device="cuda")
    ctx.gbuffer_dev(ds, cam, width, height, gb.data_ptr())
    m = Mcpg(ctx, mcpg_config(mc_tan_alpha_half=0.003, mc_distribution_dimension=1.5, lc_tan_alpha_half=0.012))
    p = pt_params(spp=spp, max_path_length=max_path_length, record_signature=False)
    for f in range(frames):
        st = m.render_dev(ds, p, cam, f, width, height, gb.data_ptr(), out.data_ptr(), stats=True)
        torch.cuda.synchronize()
        print(json.dumps({"frame": f, "mcpg_ms": round(st.ms_total, 3), "trace_ms": round(st.ms_trace, 3),
                          "shade_ms": round(st.ms_shade, 3), "segments": st.continuation_segments}), flush=True)
    st = ctx.pt_render_dev(ds, p, cam, frames, width, height, gb.data_ptr(), out.data_ptr(), stats=True)
    torch.cuda.synchronize()
    print(json.dumps({"plain_ms": round(st.ms_total, 3), "segments": st.continuation_segments}), flush=True)
    m.close()
    ds.destroy()
    ctx.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m merian_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    gd = sub.add_parser("guided")
    gd.add_argument("--width", type=int, default=480)
    gd.add_argument("--height", type=int, default=270)
    gd.add_argument("--spp", type=int, default=64)
    gd.add_argument("--train-frames", type=int, default=4)
    gd.add_argument("--frames", type=int, default=4)
    gd.add_argument("--guided-bounces", type=int, default=4)
    mf = sub.add_parser("mfma")
    mf.add_argument("dir")
    g = sub.add_parser("mcpg")
    g.add_argument("--width", type=int, default=1920)
    g.add_argument("--height", type=int, default=1080)
    g.add_argument("--spp", type=int, default=4)
    g.add_argument("--frames", type=int, default=3)
    g.add_argument("--max-path-length", type=int, default=8)
    ga = sub.add_parser("gather")
    ga.add_argument("--mib", type=int, default=4096)
    ga.add_argument("--rows", type=int, default=1 << 24)
    ga.add_argument("--reps", type=int, default=2)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    c.add_argument("--reps", type=int, default=4)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    p = sub.add_parser("pmc")
    p.add_argument("dir")
    p.add_argument("counter")
    t = sub.add_parser("traffic")
    for a in ("fetch_dir", "write_dir", "copy_fetch_dir", "copy_write_dir", "kernel"):
        t.add_argument(a)
    t.add_argument("--copy-mib", type=int, default=1024)
    t.add_argument("--gather-fetch-dir", default="", help="FETCH_SIZE run of the gather yardstick")
    t.add_argument("--gather-rows", type=int, default=1 << 24, help="rows (128-byte lines) the gather read")
    t.add_argument("--workload", default="")
    t.add_argument("--tuning-file", default="", help="JSON file whose 'tuning' object the profile ran with")
    for a in ("--stats-dir", "--valu-dir", "--l2-dir", "--dram-dir"):
        t.add_argument(a, default="")
    t.add_argument("--command", action="append", default=[], help="a command of the profile (repeatable)")
    dr = sub.add_parser("dram")
    dr.add_argument("dir")
    v = sub.add_parser("valu")
    v.add_argument("dir")
    l = sub.add_parser("l2")
    l.add_argument("dir")
    l1p = sub.add_parser("l1")
    l1p.add_argument("dir")
    args = ap.parse_args(argv)
    if args.cmd == "copy":
        copy_workload(args.mib, args.reps)
    elif args.cmd == "gather":
        gather_workload(args.mib, args.rows, args.reps)
    elif args.cmd == "mcpg":
        mcpg_workload(args.width, args.height, args.spp, args.frames, args.max_path_length)
    elif args.cmd == "guided":
        guided_workload(args.width, args.height, args.spp, args.train_frames, args.frames, args.guided_bounces)
    elif args.cmd == "traffic":
        print(json.dumps(traffic_record(args), indent=1))
    elif args.cmd in ("valu", "l2", "l1", "mfma", "dram"):
        rows = {"valu": valu, "l2": l2, "l1": l1, "mfma": mfma, "dram": dram}[args.cmd](args.dir)
        keys = list(next(iter(rows.values())).keys()) if rows else []
        print("| kernel | " + " | ".join(keys) + " |\n|---|" + "---:|" * len(keys))
        for k, r in sorted(rows.items(), key=lambda kv: -kv[1]["dispatches"] * kv[1].get("avg_ms", 1)):
            print(f"| {k} | " + " | ".join(str(r[x]) for x in keys) + " |")
    elif args.cmd == "stats":
        print("| kernel | calls | total ms | avg ms | % |\n|---|---:|---:|---:|---:|")
        for r in kernel_stats(args.dir):
            print(f"| {short_name(r['name'])} | {r['calls']} | {r['total_ns'] / 1e6:.3f} | "
                  f"{r['avg_ns'] / 1e6:.4f} | {r['pct']:.2f} |")
    else:
        print(f"| kernel | dispatches | mean {args.counter} per dispatch |\n|---|---:|---:|")
        for k, v in sorted(counter_per_kernel(args.dir, args.counter).items(), key=lambda kv: -kv[1]["mean"]):
            print(f"| {short_name(k)} | {v['dispatches']} | {v['mean']:.1f} |")
    return 0


if __name__ == "__main__":
    sys.exit(main())
