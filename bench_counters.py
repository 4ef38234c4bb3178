#!/usr/bin/env python3
"""Hardware counters of the hop kernel and of the surrogate's MFMA kernel (rocprofv3).

  python3 bench_counters.py collect --out gpurun_out/ctr     (on the GPU host)
  python3 bench_counters.py summarize --out gpurun_out/ctr   (anywhere, on the collected CSV)

collect runs bench.py once per counter group under `rocprofv3 --pmc` (one group per run: the
profiler does not split groups over passes), a plain 1 GiB device copy under FETCH_SIZE and
WRITE_SIZE (the calibration of the byte counters), and bench.py once under
`--kernel-trace --stats` for the kernels' own average durations.  Every run is a child process
with its own time limit; the first failure ends the collection.  The bench runs are one 10^4-hop
step after one warm-up step of film_625k_sites (1M excitons, 16 launches of 625 hops), the
surrogate leg included.

summarize writes profiles/hop_kernel_counters.json: per hop-kernel launch the HBM bytes read and
written (FETCH_SIZE and WRITE_SIZE scaled so that the copy reads and writes its own size), the
L2 requests, the VALU/SALU/LDS instruction counts, and for k_mlp_infer the BF16 MFMA operations
and busy cycles; each record carries build.kernel_source_hash() of the sources measured, and
bench.py uses a record only while the sources hash the same.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HOP_KERNEL = "k_hop<true, true, false>"
MLP_KERNEL = "k_mlp_infer"
COPY_KERNEL = "copyBuffer"
COPY_BYTES = 1 << 30
SIMDS = 1024                  # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9              # MI355X peak engine clock
VALU_CYCLES = 4               # a wave64 VALU instruction occupies its SIMD four cycles
BENCH_ARGS = ["--no-cpu", "--no-exact", "--no-ensemble", "--no-contact", "--steps", "1",
              "--warmup", "1"]
TRACE_ARGS = ["--no-cpu", "--no-exact", "--no-ensemble", "--no-contact", "--steps", "10",
              "--warmup", "1"]

# (name, counters, workload: "bench" or "copy"); at most 8 SQ, 4 TCC (FETCH_SIZE takes 3,
# WRITE_SIZE 2) and 2 GRBM counters per run
PASSES = [
    ("copy_fetch", ["FETCH_SIZE"], "copy"),
    ("copy_write", ["WRITE_SIZE"], "copy"),
    ("fetch", ["FETCH_SIZE"], "bench"),
    ("write", ["WRITE_SIZE"], "bench"),
    ("l2", ["TCC_HIT_sum", "TCC_MISS_sum", "TCC_REQ_sum", "TCC_EA0_RDREQ_sum"], "bench"),
    ("sq_a", ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD",
              "SQ_INSTS_VMEM_WR", "SQ_INSTS_LDS", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES",
              "GRBM_GUI_ACTIVE", "GRBM_COUNT"], "bench"),
    ("sq_b", ["SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_ANY", "SQ_WAIT_ANY",
              "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_SMEM"], "bench"),
    ("trace", [], "trace"),
    ("mfma", ["SQ_INSTS_VALU_MFMA_MOPS_BF16", "SQ_VALU_MFMA_BUSY_CYCLES", "SQ_INSTS_VALU",
              "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE"], "bench"),
]


def log(*a):
    print("[counters]", *a, flush=True)


def _bench_cmd(args):
    return [sys.executable, "-u", str(ROOT / "bench.py"), *args]


def collect(out: Path, limit: int) -> int:
    from cnt_film_monte_carlo_amd import build as B

    out.mkdir(parents=True, exist_ok=True)
    (out / "source_hash.txt").write_text(B.kernel_source_hash() + "\n")
    env = dict(os.environ, TMPDIR="/tmp")
    listing = subprocess.run(["rocprofv3", "--list-avail"], cwd="/tmp", env=env, text=True,
                             capture_output=True, timeout=120)
    (out / "list_avail.txt").write_text(listing.stdout + listing.stderr)
    avail = listing.stdout + listing.stderr if listing.returncode == 0 else None
    for name, counters, what in PASSES:
        base = [c[:-4] if c.endswith("_sum") else c for c in counters]
        if avail is not None and any(b not in avail for b in base
                                     if b not in ("FETCH_SIZE", "WRITE_SIZE")):
            log(f"pass {name}: skipped, counters not listed: "
                f"{[b for b in base if b not in avail]}")
            continue
        if what == "trace":
            prof = ["--kernel-trace", "--stats"]
            prog = _bench_cmd(TRACE_ARGS)
        else:
            prof = ["--pmc", *counters]
            prog = (_bench_cmd(["--calibrate-copy", str(COPY_BYTES)]) if what == "copy"
                    else _bench_cmd(BENCH_ARGS))
        cmd = ["timeout", "-s", "KILL", str(limit), "rocprofv3", *prof,
               "-d", str(out / name), "-o", "run", "--output-format", "csv", "--", *prog]
        log(f"pass {name}: {' '.join(prof)}")
        t0 = time.time()
        with open(out / f"{name}.stdout", "w") as fo, open(out / f"{name}.log", "w") as fe:
            rc = subprocess.run(cmd, cwd="/tmp", env=env, stdout=fo, stderr=fe).returncode
        log(f"pass {name}: exit {rc} after {time.time() - t0:.1f} s")
        if rc != 0:
            return rc
    return 0


# --------------------------------------------------------------------------------------------


def _rows(d: Path):
    for f in glob.glob(str(d / "**" / "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            yield from csv.DictReader(fh)


def per_dispatch(d: Path, kernel: str, largest_grid: bool = False) -> dict[str, float]:
    """Mean per dispatch of each counter over the dispatches of `kernel` (with largest_grid,
    only those of the largest grid)."""
    rows = [r for r in _rows(d) if kernel in r.get("Kernel_Name", "")]
    if largest_grid and rows:
        g = max(int(r["Grid_Size"]) for r in rows)
        rows = [r for r in rows if int(r["Grid_Size"]) == g]
    acc: dict[str, list[float]] = {}
    for r in rows:
        acc.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def bench_line(path: Path) -> dict:
    for line in path.read_text().splitlines():
        if line.startswith("{"):
            return json.loads(line)
    raise RuntimeError(f"no JSON line in {path}")


def trace_stats(d: Path) -> dict[str, dict]:
    out = {}
    for f in glob.glob(str(d / "**" / "*kernel_stats.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                out[r["Name"]] = {"calls": int(r["Calls"]),
                                  "avg_ms": float(r["AverageNs"]) * 1e-6,
                                  "total_ms": float(r["TotalDurationNs"]) * 1e-6}
    return out


def trace_largest_grid(d: Path, kernel: str) -> dict | None:
    """Mean duration of the dispatches of `kernel` with the largest grid in the kernel trace
    (the surrogate's 4M-pair roofline batch, not the table build's smaller calls)."""
    rows = []
    for f in glob.glob(str(d / "**" / "*kernel_trace.csv"), recursive=True):
        with open(f) as fh:
            rows += [r for r in csv.DictReader(fh) if kernel in r.get("Kernel_Name", "")]
    if not rows:
        return None
    g = max(int(r["Grid_Size_X"]) for r in rows)
    ms = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6 for r in rows
          if int(r["Grid_Size_X"]) == g]
    return {"calls": len(ms), "avg_ms": sum(ms) / len(ms), "grid": g}


def _stat(stats: dict, key: str) -> dict | None:
    hits = [v for k, v in stats.items() if key in k]
    return hits[0] if len(hits) == 1 else None


def summarize(out: Path) -> dict:
    src_hash = (out / "source_hash.txt").read_text().strip()
    cf = per_dispatch(out / "copy_fetch", COPY_KERNEL)["FETCH_SIZE"]
    cw = per_dispatch(out / "copy_write", COPY_KERNEL)["WRITE_SIZE"]
    # FETCH_SIZE and WRITE_SIZE are in KiB; the copy reads and writes COPY_BYTES each
    fetch_scale = COPY_BYTES / (cf * 1024.0)
    write_scale = COPY_BYTES / (cw * 1024.0)
    bl = bench_line(out / "fetch.stdout")
    workload = bl["config"]["workload"]
    hops = float(bl["roofline"]["hops_per_launch"])
    f = per_dispatch(out / "fetch", HOP_KERNEL)["FETCH_SIZE"] * 1024.0 * fetch_scale
    w = per_dispatch(out / "write", HOP_KERNEL)["WRITE_SIZE"] * 1024.0 * write_scale
    l2 = per_dispatch(out / "l2", HOP_KERNEL)
    sa = per_dispatch(out / "sq_a", HOP_KERNEL)
    sb = per_dispatch(out / "sq_b", HOP_KERNEL)
    stats = trace_stats(out / "trace")
    hop_t = _stat(stats, HOP_KERNEL)
    mlp_t = trace_largest_grid(out / "trace", MLP_KERNEL)
    valu = sa["SQ_INSTS_VALU"]
    rec = {
        "source_hash": src_hash,
        "kernel": HOP_KERNEL,
        "hops_per_launch": hops,
        "fetch_bytes": f,
        "write_bytes": w,
        "bytes_per_launch": f + w,
        "bytes_per_hop": (f + w) / hops,
        "l2_requests": l2.get("TCC_REQ_sum"),
        "l2_requests_per_hop": l2["TCC_REQ_sum"] / hops if "TCC_REQ_sum" in l2 else None,
        "l2_hit": l2.get("TCC_HIT_sum"),
        "l2_miss": l2.get("TCC_MISS_sum"),
        "ea_read_requests": l2.get("TCC_EA0_RDREQ_sum"),
        "valu_insts": valu,
        "valu_insts_per_wave_hop": valu / (hops / 64.0),
        "salu_insts": sa.get("SQ_INSTS_SALU"),
        "lds_insts": sa.get("SQ_INSTS_LDS"),
        "vmem_rd_insts": sa.get("SQ_INSTS_VMEM_RD"),
        "vmem_wr_insts": sa.get("SQ_INSTS_VMEM_WR"),
        "waves": sa.get("SQ_WAVES"),
        "wave_cycles": sa.get("SQ_WAVE_CYCLES"),
        "busy_cycles": sa.get("SQ_BUSY_CYCLES"),
        "gui_active": sa.get("GRBM_GUI_ACTIVE"),
        "wait_inst_any": sb.get("SQ_WAIT_INST_ANY"),
        "wait_any": sb.get("SQ_WAIT_ANY"),
        "active_inst_valu": sb.get("SQ_ACTIVE_INST_VALU"),
        "active_inst_any": sb.get("SQ_ACTIVE_INST_ANY"),
        "fma_f64_insts": sb.get("SQ_INSTS_VALU_FMA_F64"),
        "smem_insts": sb.get("SQ_INSTS_SMEM"),
        "trace_avg_launch_ms": hop_t["avg_ms"] if hop_t else None,
        "trace_launches": hop_t["calls"] if hop_t else None,
    }
    if hop_t:
        rec["valu_issue_frac"] = valu * VALU_CYCLES / (SIMDS * CLOCK_HZ * hop_t["avg_ms"] * 1e-3)
    mlp = None
    if (out / "mfma").exists():
        m = per_dispatch(out / "mfma", MLP_KERNEL, largest_grid=True)
        mops = m.get("SQ_INSTS_VALU_MFMA_MOPS_BF16")
        mlp = {
            "source_hash": src_hash,
            "kernel": MLP_KERNEL,
            "pairs_per_launch": bl["roofline_surrogate"]["pairs_per_launch"],
            "mfma_mops_bf16": mops,
            "mfma_flop": 512.0 * mops if mops is not None else None,
            "mfma_busy_cycles": m.get("SQ_VALU_MFMA_BUSY_CYCLES"),
            "valu_insts": m.get("SQ_INSTS_VALU"),
            "busy_cycles": m.get("SQ_BUSY_CYCLES"),
            "gui_active": m.get("GRBM_GUI_ACTIVE"),
            "trace_avg_launch_ms": None,
        }
        if mlp_t:
            mlp["trace_avg_launch_ms"] = mlp_t["avg_ms"]
            mlp["trace_launches"] = mlp_t["calls"]
        if mops is not None and mlp_t:
            mlp["mfma_tflops"] = 512.0 * mops / (mlp_t["avg_ms"] * 1e-3) / 1e12
    res = {
        "calibration": {"copy_bytes": COPY_BYTES, "copy_FETCH_SIZE_KiB": cf,
                        "copy_WRITE_SIZE_KiB": cw, "fetch_scale": fetch_scale,
                        "write_scale": write_scale},
        "constants": {"simds": SIMDS, "clock_hz": CLOCK_HZ, "valu_cycles": VALU_CYCLES},
        workload: rec,
        "surrogate": mlp,
        "kernel_stats": {k: v for k, v in stats.items()
                         if v["total_ms"] > 0.01 * max(s["total_ms"] for s in stats.values())},
    }
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["collect", "summarize"])
    ap.add_argument("--out", default=str(ROOT / "gpurun_out" / "ctr"))
    ap.add_argument("--limit", type=int, default=150, help="seconds per profiler run")
    ap.add_argument("--write", default=str(ROOT / "profiles" / "hop_kernel_counters.json"))
    a = ap.parse_args()
    out = Path(a.out).resolve()
    if a.mode == "collect":
        sys.exit(collect(out, a.limit))
    res = summarize(out)
    Path(a.write).write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
