"""rocprofv3 helpers: a copy workload of known size (the yardstick for the FETCH_SIZE and
WRITE_SIZE counters) and summaries of rocprofv3's CSV output.

    python -m merian_amd.profiling copy --mib 1024 --reps 4
        one plain device-to-device copy of `mib` MiB per rep (run under rocprofv3 --pmc)
    python -m merian_amd.profiling stats DIR
        per-kernel table from DIR's *_kernel_stats.csv (rocprofv3 --kernel-trace --stats)
    python -m merian_amd.profiling pmc DIR COUNTER
        per-kernel mean of COUNTER per dispatch from DIR's *_counter_collection.csv
    python -m merian_amd.profiling traffic FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR KERNEL
        calibrated HBM bytes per dispatch of KERNEL (JSON)
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
            copy_bytes: int) -> dict:
    """HBM bytes per dispatch of the kernels whose short name starts with `kernel`, from a
    FETCH_SIZE run and a WRITE_SIZE run, each scaled by what the same counter reads for a plain
    copy of `copy_bytes` bytes (which reads and writes exactly that many)."""
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

    f_kib, nf = sel(fetch_dir, "FETCH_SIZE")
    w_kib, nw = sel(write_dir, "WRITE_SIZE")
    cf = copy_bytes / (copy_kib(copy_fetch_dir, "FETCH_SIZE") * 1024.0)
    cw = copy_bytes / (copy_kib(copy_write_dir, "WRITE_SIZE") * 1024.0)
    return {"kernel": kernel, "dispatches": [nf, nw], "fetch_kib_per_dispatch": round(f_kib, 1),
            "write_kib_per_dispatch": round(w_kib, 1), "copy_bytes": copy_bytes,
            "copy_factor_fetch": round(cf, 4), "copy_factor_write": round(cw, 4),
            "traffic_bytes_per_dispatch": round(f_kib * 1024.0 * cf + w_kib * 1024.0 * cw)}


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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m merian_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
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
    args = ap.parse_args(argv)
    if args.cmd == "copy":
        copy_workload(args.mib, args.reps)
    elif args.cmd == "traffic":
        print(json.dumps(traffic(args.fetch_dir, args.write_dir, args.copy_fetch_dir, args.copy_write_dir,
                                 args.kernel, args.copy_mib << 20), indent=1))
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
