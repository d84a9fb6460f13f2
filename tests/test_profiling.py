"""The rocprofv3 summaries (merian_amd/profiling.py) and bench.py's roofline on synthetic
rocprofv3 CSV output: the calibrated traffic per dispatch, the L2 hit rate, the DRAM-request
share, and the rule that a profile counts only for the workload and tuning it was taken with."""
import csv
import json
import sys
from pathlib import Path

import pytest

from merian_amd import profiling as P

ROOT = Path(__file__).resolve().parent.parent
K = "void mrt::k_trace<false, mrt::SrcQueue, mrt::SinkQueue, false, 8>(mrt::DeviceScene, int) [clone .kd]"
COPY = "void at::native::copy_kernel(int)"


def _counters(d: Path, rows):
    d.mkdir(parents=True)
    with open(d / "x_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp",
                               "End_Timestamp"])
        w.writeheader()
        for disp, kernel, counter, value, ns in rows:
            w.writerow({"Dispatch_Id": disp, "Kernel_Name": kernel, "Counter_Name": counter, "Counter_Value": value,
                        "Start_Timestamp": 1000, "End_Timestamp": 1000 + ns})


def _stats(d: Path, rows):
    d.mkdir(parents=True)
    with open(d / "x_kernel_stats.csv", "w", newline="") as f:
        w = csv.DictWriter(f, ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"])
        w.writeheader()
        for name, calls, avg in rows:
            w.writerow({"Name": name, "Calls": calls, "TotalDurationNs": calls * avg, "AverageNs": avg,
                        "Percentage": 50.0})


def test_short_name_keeps_template_arguments():
    assert P.short_name(K) == "k_trace<false, SrcQueue, SinkQueue, false, 8>"


@pytest.fixture
def prof(tmp_path):
    # two dispatches of the kernel: 3 and 5 GiB fetched (FETCH_SIZE reads half the bytes a
    # plain copy moves on this device: the 1 GiB copy reads 512 MiB), 1 GiB written
    _counters(tmp_path / "fetch", [(1, K, "FETCH_SIZE", 3 << 20, 10_000_000), (2, K, "FETCH_SIZE", 5 << 20, 10_000_000),
                                   (3, COPY, "FETCH_SIZE", 1 << 19, 1_000_000)])
    _counters(tmp_path / "write", [(1, K, "WRITE_SIZE", 1 << 20, 10_000_000), (2, K, "WRITE_SIZE", 1 << 20, 10_000_000)])
    _counters(tmp_path / "cfetch", [(1, COPY, "FETCH_SIZE", 1 << 19, 1_000_000)])
    _counters(tmp_path / "cwrite", [(1, COPY, "WRITE_SIZE", 1 << 20, 1_000_000)])
    _counters(tmp_path / "tcc", [(1, K, "TCC_HIT_sum", 300, 10), (1, K, "TCC_MISS_sum", 100, 10)])
    _counters(tmp_path / "dram", [(1, K, "TCC_EA0_RDREQ_sum", 1000, 10), (1, K, "TCC_EA0_RDREQ_DRAM_sum", 990, 10),
                                  (1, K, "TCC_EA0_WRREQ_sum", 200, 10), (1, K, "TCC_EA0_WRREQ_DRAM_sum", 200, 10)])
    _stats(tmp_path / "stats", [(K, 2, 10_000_000), (COPY, 1, 1_000_000)])
    tuning = tmp_path / "tuning.json"
    tuning.write_text(json.dumps({"tuning": {"bvh_width": 8, "max_paths": 1 << 28}}))
    return tmp_path, tuning


def test_traffic_is_copy_calibrated(prof):
    d, _ = prof
    t = P.traffic(str(d / "fetch"), str(d / "write"), str(d / "cfetch"), str(d / "cwrite"), "k_trace<false", 1 << 30)
    assert t["dispatches"] == [2, 2]
    assert t["copy_factor_fetch"] == 2.0 and t["copy_factor_write"] == 1.0
    # mean 4 GiB read as FETCH_SIZE -> 8 GiB moved, plus 1 GiB written
    assert t["traffic_bytes_per_dispatch"] == 9 << 30


def test_traffic_reads_scaled_by_the_gather_yardstick(prof, tmp_path):
    # the gather read 2^20 whole lines (128 MiB) and FETCH_SIZE shows exactly that: reads scale
    # by 1 (scattered line fills), the upper bound by the copy's factor 2
    d, _ = prof
    G = "void at::native::vectorized_gather_kernel<16, long>(char*, char*, long)"
    _counters(tmp_path / "gfetch", [(1, G, "FETCH_SIZE", 1 << 17, 1_000_000)])
    t = P.traffic(str(d / "fetch"), str(d / "write"), str(d / "cfetch"), str(d / "cwrite"), "k_trace<false", 1 << 30,
                  str(tmp_path / "gfetch"), (1 << 20) * 128)
    assert t["gather_factor_fetch"] == 1.0 and t["copy_factor_fetch"] == 2.0
    assert t["traffic_bytes_per_dispatch"] == (4 << 30) + (1 << 30)
    assert t["traffic_upper_bytes_per_dispatch"] == (8 << 30) + (1 << 30)


def test_l2_and_dram_summaries(prof):
    d, _ = prof
    (k, v), = P.l2(str(d / "tcc")).items()
    assert k.startswith("k_trace<") and v["l2_hit_pct"] == 75.0
    (k, v), = P.dram(str(d / "dram")).items()
    assert v["rd_to_dram_pct"] == 99.0 and v["wr_to_dram_pct"] == 100.0 and v["ea_rdreq_per_dispatch"] == 1000


def test_traffic_record_carries_workload_tuning_and_counters(prof, capsys):
    d, tuning = prof
    argv = ["traffic", str(d / "fetch"), str(d / "write"), str(d / "cfetch"), str(d / "cwrite"), "k_trace<false",
            "--workload", "w", "--tuning-file", str(tuning), "--stats-dir", str(d / "stats"), "--l2-dir",
            str(d / "tcc"), "--dram-dir", str(d / "dram"), "--command", "c1", "--command", "c2"]
    assert P.main(argv) == 0
    r = json.loads(capsys.readouterr().out)
    assert r["workload"] == "w" and r["tuning"] == {"bvh_width": 8, "max_paths": 1 << 28}
    assert r["avg_ms_kernel_trace"] == 10.0 and r["l2_hit_pct"] == 75.0 and r["rd_to_dram_pct"] == 99.0
    assert r["hbm_gbs_at_kernel_trace_avg"] == round((9 << 30) / 1e-2 / 1e9, 1)
    assert r["commands"] == ["c1", "c2"]


def test_bench_roofline_uses_the_profile_only_for_its_tuning(prof, tmp_path):
    sys.path.insert(0, str(ROOT))
    import bench
    d, tuning = prof
    prof_json = tmp_path / "t.json"
    prof_json.write_text(json.dumps({"traffic_bytes_per_dispatch": 4e9, "workload": "w",
                                     "tuning": {"bvh_width": 8, "max_paths": 1 << 28}, "l2_hit_pct": 75.0}))
    r = bench.roofline("k", 1.0, 10, 8e9, str(prof_json), match={"tuning": {"bvh_width": 8, "max_paths": 1 << 28,
                                                                           "refill_lanes": 4}, "workload": "w"})
    # measured bytes over the launch time: 4 GB / 1 ms = 4 TB/s, half the peak
    assert r["achieved"] == 4000.0 and r["frac"] == 0.5 and r["traffic"] == 4e9 and r["l2_hit_pct"] == 75.0
    assert r["algorithmic_gbs"] == 8000.0 and r["traffic_over_algorithmic"] == 0.5
    for bad in ({"tuning": {"bvh_width": 4, "max_paths": 1 << 28}, "workload": "w"},
                {"tuning": {"bvh_width": 8, "max_paths": 1 << 28}, "workload": "other"}):
        r = bench.roofline("k", 1.0, 10, 8e9, str(prof_json), match=bad)
        assert r["traffic"] is None and r["frac"] == 1.0 and "algorithmic" in r["basis"]


def test_l1_request_summary(tmp_path):
    # one dispatch of 10 ms on a 256-CU, 8-XCD agent busy for 24 M cycles per XCD: 2 G L1->L2
    # reads of 3 G accesses; the addresser busy 18 M cycles per CU and stalled 12 M per CU
    _counters(tmp_path / "l1", [(1, K, "TCP_TCC_READ_REQ_sum", 2e9, 10_000_000),
                                (1, K, "TCP_TOTAL_CACHE_ACCESSES_sum", 3e9, 10_000_000),
                                (1, K, "TA_BUSY_avr", 18e6, 10_000_000),
                                (1, K, "TA_ADDR_STALLED_BY_TC_CYCLES_sum", 256 * 12e6, 10_000_000),
                                (1, K, "GRBM_GUI_ACTIVE", 8 * 24e6, 10_000_000)])
    with open(tmp_path / "l1" / "x_agent_info.csv", "w", newline="") as f:
        w = csv.DictWriter(f, ["Agent_Type", "Name", "Cu_Count", "Simd_Count", "Num_Xcc", "Max_Engine_Clk_Fcompute"])
        w.writeheader()
        w.writerow({"Agent_Type": "CPU", "Name": "EPYC", "Cu_Count": 16, "Simd_Count": 0, "Num_Xcc": 0,
                    "Max_Engine_Clk_Fcompute": 0})
        w.writerow({"Agent_Type": "GPU", "Name": "gfx950", "Cu_Count": 256, "Simd_Count": 1024, "Num_Xcc": 8,
                    "Max_Engine_Clk_Fcompute": 2400})
    (k, v), = P.l1(str(tmp_path / "l1")).items()
    assert k.startswith("k_trace<") and v["l2_read_req_per_dispatch"] == 2e9 and v["l2_read_greq_per_s"] == 200.0
    assert v["l1_hit_pct"] == 33.3 and v["ta_busy_pct"] == 75.0 and v["ta_stalled_on_l1_pct"] == 50.0


def test_bench_roofline_ignores_a_profile_of_another_kernel(tmp_path):
    sys.path.insert(0, str(ROOT))
    import bench
    prof_json = tmp_path / "t.json"
    prof_json.write_text(json.dumps({"traffic_bytes_per_dispatch": 4e9, "workload": "w", "kernel": "k_trace<..., 8>",
                                     "tuning": {"bvh_width": 8}}))
    m = {"tuning": {"bvh_width": 8, "bvh8_order": 1}, "workload": "w"}
    assert bench.roofline("k_trace<..., 8>", 1.0, 10, 8e9, str(prof_json), match=m)["traffic"] == 4e9
    assert bench.roofline("k_trace<..., 9>", 1.0, 10, 8e9, str(prof_json), match=m)["traffic"] is None
    assert bench.trace_variant({"bvh_width": 8, "bvh8_order": 1}) == 9 and bench.trace_variant({"bvh_width": 8}) == 8
    assert bench.trace_variant({"bvh_width": 4, "bvh8_order": 1}) == 4
