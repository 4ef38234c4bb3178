"""The measurement plumbing on the CPU: bench_counters.py's summary of rocprofv3 counter CSVs
(per-dispatch means of the hop kernel, the copy calibration of FETCH_SIZE / WRITE_SIZE, the
derived per-hop figures) on synthetic files, and bench.py's refusal of a counter record taken on
other kernel sources or another launch size."""
import csv
import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

HOP = "void cntkmc::k_hop<true, true, false>(cntkmc::hop_args)"
MLP = "(anonymous namespace)::k_mlp_infer((anonymous namespace)::infer_args)"
COPY = "__amd_rocclr_copyBuffer"


def _csv(d: Path, rows):
    d.mkdir(parents=True)
    with open(d / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for r in rows:
            w.writerow(r)


def _fake_collection(out: Path, src_hash: str):
    out.mkdir(parents=True, exist_ok=True)
    (out / "source_hash.txt").write_text(src_hash + "\n")
    _csv(out / "copy_fetch", [(1, 1, COPY, "FETCH_SIZE", 524288.0)])
    _csv(out / "copy_write", [(1, 1, COPY, "WRITE_SIZE", 1048576.0)])
    bench = {"config": {"workload": "film_625k_sites"},
             "roofline": {"hops_per_launch": 1000.0},
             "roofline_surrogate": {"pairs_per_launch": 64}}
    for name in ("fetch", "write", "l2", "sq_a", "sq_b", "mfma"):
        (out / f"{name}.stdout").write_text("progress\n" + json.dumps(bench) + "\n")
    # two dispatches each: the summary takes their mean
    _csv(out / "fetch", [(1, 1, HOP, "FETCH_SIZE", 1.0), (2, 1, HOP, "FETCH_SIZE", 3.0),
                         (3, 1, "other", "FETCH_SIZE", 1e9)])
    _csv(out / "write", [(1, 1, HOP, "WRITE_SIZE", 4.0), (2, 1, HOP, "WRITE_SIZE", 4.0)])
    _csv(out / "l2", [(1, 1, HOP, "TCC_REQ_sum", 500.0), (1, 1, HOP, "TCC_HIT_sum", 400.0),
                      (1, 1, HOP, "TCC_MISS_sum", 100.0), (1, 1, HOP, "TCC_EA0_RDREQ_sum", 8.0)])
    _csv(out / "sq_a", [(1, 1, HOP, "SQ_INSTS_VALU", 3200.0), (1, 1, HOP, "SQ_WAVES", 16.0)])
    _csv(out / "sq_b", [(1, 1, HOP, "SQ_WAIT_ANY", 10.0)])
    _csv(out / "mfma", [(1, 2, MLP, "SQ_INSTS_VALU_MFMA_MOPS_BF16", 1.0),
                        (2, 8, MLP, "SQ_INSTS_VALU_MFMA_MOPS_BF16", 1000.0)])
    (out / "trace").mkdir()
    with open(out / "trace" / "run_kernel_stats.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs"])
        w.writerow([HOP, 10, 20_000_000, 2_000_000])
        w.writerow([MLP, 3, 3_000_000, 1_000_000])
    # the inference kernel's own time: the dispatches of the largest grid only (3 ms and 1 ms,
    # not the 0.5 ms of a smaller call)
    with open(out / "trace" / "run_kernel_trace.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X"])
        w.writerow([MLP, 0, 500_000, 768])
        w.writerow([MLP, 1_000_000, 4_000_000, 196_608])
        w.writerow([MLP, 5_000_000, 6_000_000, 196_608])


def test_summary_of_synthetic_counters(tmp_path):
    import bench_counters as BC
    out = tmp_path / "ctr"
    _fake_collection(out, "abc123")
    res = BC.summarize(out)
    cal = res["calibration"]
    assert cal["fetch_scale"] == pytest.approx(2.0) and cal["write_scale"] == pytest.approx(1.0)
    r = res["film_625k_sites"]
    assert r["source_hash"] == "abc123"
    assert r["fetch_bytes"] == pytest.approx(2.0 * 1024 * 2.0)      # mean 2 KiB, scaled x 2
    assert r["write_bytes"] == pytest.approx(4.0 * 1024)
    assert r["bytes_per_hop"] == pytest.approx((4096 + 4096) / 1000)
    assert r["l2_requests_per_hop"] == pytest.approx(0.5)
    assert r["valu_insts_per_wave_hop"] == pytest.approx(3200 / (1000 / 64))
    assert r["trace_avg_launch_ms"] == pytest.approx(2.0)
    assert r["valu_issue_frac"] == pytest.approx(3200 * 4 / (1024 * 2.4e9 * 2e-3))
    m = res["surrogate"]
    assert m["mfma_mops_bf16"] == 1000.0                              # the largest grid only
    assert m["mfma_flop"] == 512_000.0
    assert m["trace_avg_launch_ms"] == pytest.approx(2.0) and m["trace_launches"] == 2
    assert m["mfma_tflops"] == pytest.approx(512_000.0 / 2e-3 / 1e12)


def test_bench_refuses_stale_counter_records(tmp_path, monkeypatch):
    import bench
    from cnt_film_monte_carlo_amd import build as B
    prof = tmp_path / "profiles"
    prof.mkdir()
    monkeypatch.setattr(bench, "ROOT", tmp_path)
    rec = {"source_hash": B.kernel_source_hash(), "hops_per_launch": 625e6,
           "bytes_per_launch": 5.0e9}
    (prof / "hop_kernel_counters.json").write_text(json.dumps({"film_625k_sites": rec}))
    assert bench.counter_record("film_625k_sites", 625e6)["bytes_per_launch"] == 5.0e9
    assert bench.counter_record("film_625k_sites", 1e6) is None        # another launch size
    rec["source_hash"] = "0" * 64
    (prof / "hop_kernel_counters.json").write_text(json.dumps({"film_625k_sites": rec}))
    assert bench.counter_record("film_625k_sites", 625e6) is None      # other kernel sources
    assert bench.counter_record("film_1600k_sites", 625e6) is None     # no record


def test_source_hash_follows_the_sources(tmp_path, monkeypatch):
    from cnt_film_monte_carlo_amd import build as B
    h0 = B.kernel_source_hash()
    assert h0 == B.kernel_source_hash() and len(h0) == 64
    csrc = tmp_path / "csrc"
    inc = tmp_path / "include"
    csrc.mkdir()
    inc.mkdir()
    for f in B.CSRC.glob("*"):
        if f.suffix in (".hip", ".h"):
            (csrc / f.name).write_bytes(f.read_bytes())
    for f in (B.ROOT / "include").glob("*.h"):
        (inc / f.name).write_bytes(f.read_bytes())
    monkeypatch.setattr(B, "CSRC", csrc)
    monkeypatch.setattr(B, "ROOT", tmp_path)
    assert B.kernel_source_hash() == h0                                 # location-independent
    k = csrc / "kmc_kernels.hip"
    k.write_bytes(k.read_bytes() + b"\n// changed\n")
    assert B.kernel_source_hash() != h0
