#!/usr/bin/env python3
"""Benchmark of the KMC exciton-hopping hot path on one MI355X.

Workload (film_625k_sites): 5,000 straight CNTs of length U[0.5, 2] um in a 5x5x0.2 um box,
10 nm segment sites (624,739 sites), chiralities (7,5)/(8,6)/(9,7), cutoff 20 nm, y periodic,
injection at x <= 0 and drain at x >= 5 um; 1M excitons at 300 K.

Rates: the nine donor/acceptor 4-D tables on the input.json grid (21 x 11 x 11 x 11 nodes),
built twice -- by the exact first-order rate on the GPU (create_davoody_scatt_table) and by
the BF16 surrogate with exact fallback for flagged nodes; both build times are reported.  The
hop run uses the surrogate-built tables.  A step is one pass of the hop kernel over every
exciton for STEP_HOPS = 10^4 hops (10^10 hops), with MSD sampling and per-site hop counts on;
the default 10 timed steps are the full 10^5 hops per exciton.  A second run on the
exact-built tables with the same seeds gives the diffusion coefficient the surrogate run is
compared against.

--workload film_1600k_sites runs the larger film instead: 80,000 CNTs of 200 nm in a
20x20x0.5 um box (1.6M sites), 16M excitons, with a 64-network ensemble leg.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "exciton hops/sec (1 GPU) at D within 2%; rate-table build s (MLP vs exact)"
STEP_HOPS = 10_000
RESORT_HOPS = 625           # hop-kernel sub-launch length between exciton sorts
HBM_PEAK = 8.0e12
BF16_PEAK = 2.5e15          # dense MFMA, MI355X
MLP_FLOP_PER_PAIR = 2 * (13 * 128 + 3 * 128 * 128 + 128)   # 13 used inputs, 4 x 128, 1 output
TEMPERATURE = 300.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def counter_record(key, hops_per_launch):
    """The record `key` of profiles/hop_kernel_counters.json if its kernel-source hash is that
    of the sources in this tree (and, for a hop workload, its hops per launch are this run's);
    else None."""
    from cnt_film_monte_carlo_amd import build as B

    f = ROOT / "profiles" / "hop_kernel_counters.json"
    if not f.exists():
        return None
    try:
        rec = json.loads(f.read_text()).get(key)
    except ValueError:
        return None
    if not rec or rec.get("source_hash") != B.kernel_source_hash():
        if rec:
            log(f"counter record '{key}' was taken on other kernel sources: not used")
        return None
    if hops_per_launch is not None and round(rec["hops_per_launch"]) != round(hops_per_launch):
        log(f"counter record '{key}' has {rec['hops_per_launch']:.4g} hops per launch, this run "
            f"{hops_per_launch:.4g}: not used")
        return None
    return rec


def cpu_model() -> str:
    """The host CPU's model string (lscpu's "Model name", else /proc/cpuinfo's)."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _sized_sample(make, run, E, seconds, h_max):
    """Hops per exciton H of a fresh sample of E excitons sized so that run(sample, H) takes
    about `seconds` (a short pilot first, one re-size if the run came out short); returns
    (H, wall seconds of the last run)."""
    pilot = make()
    run(pilot, 5)
    h = 20
    while True:
        t0 = time.perf_counter()
        run(pilot, h)
        dt = time.perf_counter() - t0
        if dt > 0.5 or h >= h_max:
            break
        h = min(h_max, h * 4)
    H = int(max(50, min(h_max, h * seconds / max(dt, 1e-6))))
    for _ in range(2):
        ex = make()
        t0 = time.perf_counter()
        run(ex, H)
        dt = time.perf_counter() - t0
        if dt >= 0.8 * seconds or H >= h_max:
            break
        H = int(max(50, min(h_max, H * 1.1 * seconds / dt)))
    return H, dt


def cpu_baseline(net, w, tables, csr, seconds=12.0):
    """The oracle (CPU restatement of the reference's OpenMP path) on bounded samples of the
    same workload, each sized to about `seconds` of CPU work: 'faithful' recomputes each site's
    neighbour list and rates on every hop as scatterer.cpp:12 does; 'cached' reads the same CSR
    as the GPU."""
    from cnt_film_monte_carlo_amd import kmc
    from oracle import oracle as orc

    th, z, a1, a2, rates = tables
    onet = orc.Net(net, w.cutoff, 0.4e-9, kmc.PERIODIC_Y, (0, 0, 0), (0, w.period_y[1], 0))
    onet.set_table(th, z, a1, a2, rates)
    threads = orc.library().orc_max_threads()
    model = cpu_model()
    E = 4096
    H, dt = _sized_sample(lambda: orc.Excitons(csr, net["flags"], seed=99, n_excitons=E),
                          lambda ex, h: ex.run(h, net=onet), E, seconds, 200_000)
    faithful = E * H / dt
    log(f"cpu faithful: {E}x{H} hops in {dt:.2f} s = {faithful:.3e} hops/s on {threads} threads "
        f"({model})")
    Ec = 4 * E
    Hc, dtc = _sized_sample(lambda: orc.Excitons(csr, net["flags"], seed=99, n_excitons=Ec),
                            lambda ex, h: ex.run(h), Ec, seconds, 2_000_000)
    cached = Ec * Hc / dtc
    log(f"cpu cached-CSR: {Ec}x{Hc} hops in {dtc:.2f} s = {cached:.3e} hops/s")
    return {
        "value": faithful,
        "unit": "hops/s",
        "cores": threads,
        "cpu_model": model,
        "kind": "port",
        "sample": (f"{E} excitons x {H} hops ({dt:.1f} s) on the same {net['n']:,}-site "
                   f"network, neighbours and rates recomputed per hop (scatterer.cpp:12), "
                   f"OpenMP over excitons"),
        "cached_csr_value": cached,
        "cached_csr_sample": (f"{Ec} excitons x {Hc} hops ({dtc:.1f} s) over the GPU-built CSR, "
                              f"OpenMP over excitons"),
    }


def rate_tables(w, exact_tables=True):
    """Nine exact and nine surrogate-built tables at 300 K on the input.json grid, timed."""
    import numpy as np

    from cnt_film_monte_carlo_amd import exciton as X
    from cnt_film_monte_carlo_amd import kmc
    from cnt_film_monte_carlo_amd import surrogate as S

    grid = kmc.default_grids()
    sur = S.RateSurrogate.load()
    cls = [S.CHIRALITIES.index(c) for c in w.chiralities]
    # the CNT exciton models: v(q) and PI(q) on the GPU, timed beside the all-host build
    t0 = time.perf_counter()
    for c in w.chiralities:
        X.CntModel(*c, 10).close()
    host_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    models = {c: X.CntModel(*c, 10, device=0) for c in w.chiralities}
    gpu_s = time.perf_counter() - t0
    log(f"CNT models {w.chiralities}: {gpu_s * 1e3:.1f} ms with the GPU sums, "
        f"{host_s * 1e3:.1f} ms on the host cores")
    cnt_s = dict(gpu=gpu_s, host=host_s, host_cores=len(os.sched_getaffinity(0)))
    trs = []
    for ci in w.chiralities:
        for cj in w.chiralities:
            t = X.Transfer(models[ci], models[cj], TEMPERATURE)
            t.set_clamp(sur.domain.box_lo, sur.domain.box_hi)
            t.rates([0.1], [2e-9], [0.0], [0.0])     # device upload outside the timings
            trs.append(t)
    sur.rates([0.1], [2e-9], [0.0], [0.0], 0, 0, TEMPERATURE)
    exact, mlp = [], []
    exact_s = mlp_s = 0.0
    fallback = 0
    k = 0
    for i in range(len(cls)):
        for j in range(len(cls)):
            tab, nfb, ms = sur.table(cls[i], cls[j], TEMPERATURE, trs[k], *grid)
            if exact_tables:
                ex = trs[k].table(*grid)
                exact_s += trs[k].last_ms / 1e3
            else:
                ex = tab
            mlp_s += ms / 1e3
            fallback += nfb
            exact.append(ex)
            mlp.append(tab)
            k += 1
    exact = np.stack(exact)
    mlp = np.stack(mlp)
    if not exact_tables:
        log(f"tables: surrogate {mlp_s * 1e3:.1f} ms, {fallback}/{mlp.size} nodes exact")
        return dict(grid=grid, exact=None, mlp=mlp, exact_s=None, mlp_s=mlp_s, fallback=fallback,
                    nodes=int(mlp.size), err_median=None, err_p99=None, err_over=None,
                    surrogate=sur, transfers=trs, models=models, cnt_s=cnt_s)
    err = S.relative_error(np.log10(np.maximum(mlp, 1e-300)), np.log10(np.maximum(exact, 1e-300)))
    log(f"tables: exact {exact_s:.3f} s, surrogate {mlp_s * 1e3:.1f} ms ({exact_s / mlp_s:.0f}x), "
        f"{fallback}/{exact.size} nodes exact, error median {np.median(err):.4f} p99 "
        f"{np.quantile(err, 0.99):.4f}")
    return dict(grid=grid, exact=exact, mlp=mlp, exact_s=exact_s, mlp_s=mlp_s, fallback=fallback,
                nodes=int(exact.size), err_median=float(np.median(err)),
                err_p99=float(np.quantile(err, 0.99)), err_over=float(np.mean(err > 0.05)),
                surrogate=sur, transfers=trs, models=models, cnt_s=cnt_s)


def exact_table_run(w, net, tabs, E, args, t_sample_max, D):
    """The timed run's hop sequence on the exact-built tables (same seeds): its diffusion
    coefficient and the surrogate run's relative difference from it."""
    from cnt_film_monte_carlo_amd import kmc
    from cnt_film_monte_carlo_amd import networks as nw

    th, z, a1, a2 = tabs["grid"]
    ctx_x = nw.make_context(w, net, seed=2024)
    ctx_x.set_rate_table(th, z, a1, a2, tabs["exact"])
    ctx_x.build_cells()
    ctx_x.build_rates(kmc.RATE_TABLE)
    ctx_x.inject(E)
    ctx_x.set_samples(args.bins, t_sample_max)
    for _ in range(args.steps):
        ctx_x.run(STEP_HOPS, count_sites=False, timed=False)
    red_x, _ = ctx_x.reduce_samples(32)
    ctx_x.close()
    try:
        Dx, cix = kmc.estimate_diffusion(red_x, t_sample_max, args.bins // 4, args.bins)
        Dx, cix = Dx.tolist(), cix.tolist()
    except kmc.KmcError:
        return None, None, None
    d_rel = None
    if D is not None:
        d_rel = [(D[k] - Dx[k]) / Dx[k] if Dx[k] else None for k in range(3)]
    log(f"exact-table run: D = {Dx}; surrogate vs exact {d_rel}")
    return Dx, cix, d_rel


def ensemble_leg(sur, n_networks=256, hops=10_000, plain_subset=16, bins=16):
    """256 networks of uneven size (film_5k_sites rules, 5k-200k sites, excitons = 0.2 x sites,
    T ~ U[150, 400] K), per-pair surrogate rates at each network's T, one hop launch per
    sub-launch for all of them (schedule from the autotuner); per-network plain runs of a subset
    with the same seeds for the comparison."""
    import numpy as np

    from cnt_film_monte_carlo_amd import exciton as X
    from cnt_film_monte_carlo_amd import kmc
    from cnt_film_monte_carlo_amd import networks as nw

    t0 = time.perf_counter()
    ws = nw.ensemble(n_networks)
    model = X.CntModel(7, 5, 10)
    nets, ctxs, trs = [], [], []
    build_ms = 0.0
    fallback = 0
    for k, w in enumerate(ws):
        net = nw.build_network(w)
        tr = X.Transfer(model, model, w.temperature)
        tr.set_clamp(sur.domain.box_lo, sur.domain.box_hi)
        ctx = nw.make_context(w, net, seed=5000 + k)
        ctx.build_cells()
        ctx.build_rates(kmc.RATE_PAIR)
        ms, nfb = ctx.build_rates_mlp(sur, [0], w.temperature, [tr])
        build_ms += ms
        fallback += nfb
        nets.append(net)
        ctxs.append(ctx)
        trs.append(tr)
    sites = sum(c.n_sites for c in ctxs)
    excitons = sum(w.n_excitons for w in ws)
    log(f"ensemble: {n_networks} networks, {sites} sites, {excitons} excitons built in "
        f"{time.perf_counter() - t0:.1f} s (surrogate pair rates {build_ms / 1e3:.2f} s, "
        f"{fallback} pairs exact)")
    # pilot: the MSD window of every network (its earliest exciton clock after the hop budget)
    for c, w in zip(ctxs, ws):
        c.inject(w.n_excitons)
    ens = kmc.Ensemble(ctxs)
    ens.run(hops, count_sites=False)
    t_s = []
    for c in ctxs:
        tn = c.get_excitons()["t_next"]
        t_s.append(0.8 * float(np.min(tn[np.isfinite(tn)])))
    ens.close()
    for c, w, ts in zip(ctxs, ws, t_s):
        c.inject(w.n_excitons)
        c.set_samples(bins, ts)
    ens = kmc.Ensemble(ctxs)
    best, rates = ens.autotune_schedule(trial_hops=1250)
    log(f"ensemble autotune: resort_hops {best}; candidates hops/s {np.round(rates / 1e9, 1)} e9")
    t1 = time.perf_counter()
    ms = ens.run(hops)
    wall = time.perf_counter() - t1
    hops_done = sum(int(c.get_excitons()["hops"].sum()) for c in ctxs)
    log(f"ensemble run: {hops_done:.3e} hops in {wall:.3f} s ({hops_done / wall:.3e} hops/s)")
    ens.close()
    # plain runs of a subset, same seeds
    sub = list(range(0, n_networks, max(1, n_networks // plain_subset)))[:plain_subset]
    plain_hops, plain_s, d_diff = 0, 0.0, 0.0
    for k in sub:
        w, ts = ws[k], t_s[k]
        c = nw.make_context(w, nets[k], seed=5000 + k)
        c.build_cells()
        c.build_rates(kmc.RATE_PAIR)
        c.build_rates_mlp(sur, [0], w.temperature, [trs[k]])
        c.inject(w.n_excitons)
        c.set_samples(bins, ts)
        c.set_hop_schedule(best)
        c.sync()
        tp = time.perf_counter()
        c.run(hops)
        c.sync()
        plain_s += time.perf_counter() - tp
        plain_hops += int(c.get_excitons()["hops"].sum())
        De, _ = kmc.estimate_diffusion(ctxs[k].reduce_samples(8)[0], ts, bins // 4, bins)
        Dp, _ = kmc.estimate_diffusion(c.reduce_samples(8)[0], ts, bins // 4, bins)
        d_diff = max(d_diff, float(np.max(np.abs(De - Dp))))
        same = np.array_equal(ctxs[k].get_excitons()["hash"], c.get_excitons()["hash"])
        if not same:
            d_diff = float("inf")
        c.close()
    log(f"plain runs of {len(sub)} networks: {plain_hops / plain_s:.3e} hops/s; "
        f"max |D_ensemble - D_plain| = {d_diff}")
    for c in ctxs:
        c.close()
    return {"networks": n_networks, "sites": sites, "excitons": excitons,
            "hops_per_exciton": hops, "hops": hops_done, "seconds": wall,
            "hops_per_s": hops_done / wall, "kernel_ms": ms, "resort_hops_autotuned": best,
            "autotune_candidates_hops_per_s": rates.tolist(),
            "rate_build_s": build_ms / 1e3, "pairs_exact_fallback": fallback,
            "plain_subset_networks": len(sub), "plain_subset_hops_per_s": plain_hops / plain_s,
            "max_abs_D_diff_vs_plain": d_diff,
            "note": "ensemble members are bit-identical to per-network runs (same seeds); "
                    "D compared on the plain-run subset"}


def contact_leg(w, net, tabs, n_seg=20, c1_pop=100_000, dwell_mult=100.0, warmup=2, steps=20):
    """The reference's second workflow on the same film: contact (steady-state) mode with the
    transport along y, n_seg sections, c1_pop particles held in the first contact and the last
    emptied each step (monte_carlo.h:443-491), from the initial linear profile.  A step is
    dwell_mult mean dwell times, so a particle makes ~dwell_mult hops per step.  Open axes: the
    contacts are the film's two y ends."""
    import numpy as np

    from cnt_film_monte_carlo_amd import kmc

    flags = np.zeros(net["n"], np.uint8)
    ctx = kmc.Context(kmc.Config(cutoff=w.cutoff, min_distance=0.4e-9, seed=4242))
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"], net["left"],
                  net["right"], net["chirality"], flags)
    th, z, a1, a2 = tabs["grid"]
    ctx.set_rate_table(th, z, a1, a2, tabs["mlp"])
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    csr = ctx.get_csr()
    deg = np.diff(csr["row_ptr"])
    gamma = csr["cum"][csr["row_ptr"][1:][deg > 0] - 1]
    dt = dwell_mult * float(np.mean(1.0 / gamma))
    capacity = 4 * n_seg * c1_pop // 2
    ctx.contact_init(n_seg, c1_pop, 0, capacity)
    ctx.contact_run(dt, warmup)
    h0, ms0 = ctx.contact_stats()
    ctx.sync()
    t0 = time.perf_counter()
    pop, cur, live = ctx.contact_run(dt, steps)
    wall = time.perf_counter() - t0
    h1, ms1 = ctx.contact_stats()
    ctx.close()
    hops = h1 - h0
    kern_s = (ms1 - ms0) / 1e3
    particle_steps = int(pop.sum())
    # algorithmic bytes: a hop reads the row header and the cumulative rates (24 + 8 d, d the
    # mean out-degree: the visit weights of contact mode are not counted) and a particle step
    # reads its slot (site, t_next, hops, id, y: 32 B) and writes it back (28 B)
    bpe = 24.0 + 8.0 * float(deg.mean())
    achieved = (hops * bpe + particle_steps * 60.0) / kern_s / 1e9
    log(f"contact mode: {steps} steps of {dt:.3e} s, {particle_steps / steps:.0f} particles, "
        f"{hops:.3e} hops in {wall:.3f} s ({hops / wall:.3e} hops/s; step kernels "
        f"{kern_s * 1e3:.1f} ms, {hops / kern_s:.3e} hops/s)")
    return {
        "sites": int(net["n"]), "n_seg": n_seg, "c1_pop": c1_pop, "c2_pop": 0,
        "dt_s": dt, "steps": steps, "warmup": warmup,
        "particles_mean": particle_steps / steps, "hops": int(hops), "seconds": wall,
        "hops_per_s": hops / wall, "step_kernel_ms": kern_s * 1e3,
        "step_kernel_hops_per_s": hops / kern_s,
        "roofline": {"kernel": "k_contact_wave", "bound": "hbm", "achieved": achieved,
                     "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                     "bytes_per_hop": bpe, "bytes_per_particle_step": 60.0},
        "note": "whole steps (particle steps + repopulation + per-step outputs) timed by the "
                "wall clock; the step kernels by HIP events on the context stream",
    }


def surrogate_roofline(sur, n=1 << 22, reps=5):
    """Fused BF16 inference kernel on n pair geometries: algorithmic FLOP / kernel time."""
    import numpy as np

    from cnt_film_monte_carlo_amd import surrogate as S

    rng = np.random.default_rng(3)
    g = S.sample_geometries(rng, n, sur.domain)
    ms = []
    for _ in range(reps):
        _, _, k = sur.rates(*g, 0, 0, TEMPERATURE)
        ms.append(k)
    avg = float(np.mean(ms[1:]))
    achieved = n * MLP_FLOP_PER_PAIR / (avg * 1e-3)
    log(f"surrogate inference: {n} pairs in {avg:.3f} ms = {achieved / 1e12:.1f} TFLOP/s")
    return {"kernel": "k_mlp_infer", "bound": "mfma", "achieved": achieved / 1e12,
            "peak": BF16_PEAK / 1e12, "unit": "TFLOP/s", "frac": achieved / BF16_PEAK,
            "traffic": None, "flop_per_pair": MLP_FLOP_PER_PAIR, "pairs_per_launch": n,
            "avg_launch_ms": avg, "pairs_per_s": n / (avg * 1e-3)}


def cpu_table_baseline(tabs, per_class=1):
    """The oracle's literal first_order (exciton_transfer.cpp:380-441, one core per node) on a
    sample of table nodes, scaled to the nine tables."""
    import numpy as np

    from oracle import oracle as orc

    models = tabs["models"]
    chir = list(models)
    th, z, a1, a2 = tabs["grid"]
    rng = np.random.default_rng(4)
    secs = []
    for ci in chir:
        td = orc.tube_from_model(models[ci])
        for cj in chir:
            ta = orc.tube_from_model(models[cj])
            for _ in range(per_class):
                t0 = time.perf_counter()
                orc.first_order(td, ta, TEMPERATURE, 4.0e-3 * 1.6e-19, float(rng.choice(z)),
                                float(rng.choice(a1)), float(rng.choice(a2)), float(rng.choice(th)))
                secs.append((ci, cj, time.perf_counter() - t0))
    per_pair = {}
    for ci, cj, s in secs:
        per_pair.setdefault((ci, cj), []).append(s)
    nodes = tabs["nodes"] // (len(chir) ** 2)
    total = sum(np.mean(v) * nodes for v in per_pair.values())
    log(f"cpu table build (oracle first_order, 1 core): {total:.0f} s estimated")
    return {"value": total, "unit": "s (nine tables)", "cores": 1, "kind": "port",
            "sample": f"{len(secs)} nodes timed, mean per chirality pair x {nodes} nodes"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="film_625k_sites",
                    choices=["film_625k_sites", "film_1600k_sites"],
                    help="film_625k_sites: 5,000 CNTs in 5x5x0.2 um, 1M excitons, 256-network "
                         "ensemble; film_1600k_sites: 80,000 CNTs in 20x20x0.5 um, 16M excitons, "
                         "64-network ensemble")
    ap.add_argument("--excitons", type=int, default=0, help="0: the workload's own count")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-exact", action="store_true",
                    help="skip the exact table build and the exact-table comparison run")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 256-network ensemble leg")
    ap.add_argument("--no-contact", action="store_true",
                    help="skip the contact (steady-state) mode leg")
    ap.add_argument("--bins", type=int, default=64)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {n * 4} bytes x 3")
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the KMC path would shard excitons across GPUs, but "
              "it runs and reports a single MI355X only (--gpus 1).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from cnt_film_monte_carlo_amd import kmc
    from cnt_film_monte_carlo_amd import networks as nw

    w = nw.film_large() if args.workload == "film_625k_sites" else nw.film_xlarge()
    n_networks = 256 if args.workload == "film_625k_sites" else 64
    t0 = time.perf_counter()
    net = nw.build_network(w)
    log(f"network: {net['n']} sites in {time.perf_counter() - t0:.2f} s")
    tabs = rate_tables(w, exact_tables=not args.no_exact)
    th, z, a1, a2 = tabs["grid"]
    tables = (th, z, a1, a2, tabs["mlp"])
    ctx = nw.make_context(w, net, seed=2024)
    ctx.set_rate_table(th, z, a1, a2, tabs["mlp"])
    ctx.set_hop_schedule(RESORT_HOPS)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    ctx.sync()
    build_s = time.perf_counter() - t0
    nnz = ctx.nnz
    log(f"CSR: {nnz} pairs built in {build_s * 1e3:.2f} ms")
    csr = ctx.get_csr()
    deg = np.diff(csr["row_ptr"])

    E = args.excitons or w.n_excitons
    # warmup (untimed) then a fresh injection for the timed region
    ctx.inject(E)
    for i in range(args.warmup):
        ms = ctx.run(STEP_HOPS)
        log(f"warmup step {i}: {ms:.1f} ms")
    # The warmup replays the first warmup*STEP_HOPS hops of the timed run (same seed), so its
    # earliest exciton clock is a time every exciton reaches: the MSD window ends there.
    t_probe = ctx.get_excitons()["t_next"]
    fin = np.isfinite(t_probe)
    t_sample_max = float(np.min(t_probe[fin])) if args.warmup and fin.any() else 1e-9
    ctx.inject(E)
    ctx.set_samples(args.bins, t_sample_max)
    ctx.sync()

    kernel_ms, hop_ms, n_launch = [], [], 0
    t_start = time.perf_counter()
    for i in range(args.steps):
        ms = ctx.run(STEP_HOPS, count_sites=True, timed=True)
        kernel_ms.append(ms)
        hop_ms.append(ctx.hop_time())
        n_launch += ctx.run_stats(live=False)[0]
        log(f"step {i}: {ms:.1f} ms, of which hop kernel {hop_ms[-1]:.1f} ms")
    red, red_ms = ctx.reduce_samples(32)
    ctx.sync()
    t_end = time.perf_counter()
    wall = t_end - t_start

    st = ctx.get_excitons()
    hops_total = int(st["hops"].sum())
    counts = ctx.get_hop_counts().astype(np.float64)
    d_v = float((counts * deg).sum() / counts.sum())
    bytes_per_hop = 24.0 + 8.0 * d_v
    launches = n_launch / args.steps                    # hop-kernel launches per step
    avg_ms = float(np.sum(hop_ms)) / n_launch           # hop-kernel device time per launch
    hops_per_launch = hops_total / n_launch
    achieved = hops_per_launch * bytes_per_hop / (avg_ms * 1e-3)
    try:
        D, ci = kmc.estimate_diffusion(red, t_sample_max, args.bins // 4, args.bins)
        D, ci = D.tolist(), ci.tolist()
    except kmc.KmcError:
        D, ci = None, None
    log(f"hops {hops_total:.3e} in {wall:.2f} s; D = {D}")
    ctx.close()

    # the same run on the exact-built tables: the diffusion coefficient to compare against
    Dx = cix = d_rel = None
    if tabs["exact"] is not None:
        Dx, cix, d_rel = exact_table_run(w, net, tabs, E, args, t_sample_max, D)
    sur_roof = surrogate_roofline(tabs["surrogate"])
    ens = None if args.no_ensemble else ensemble_leg(tabs["surrogate"], n_networks=n_networks)
    contact = None if args.no_contact else contact_leg(w, net, tabs)


    # HBM bytes, L2 requests and VALU instructions per hop-kernel launch from the rocprofv3
    # counter passes of bench_counters.py, used only if they were taken on these sources
    ctr = counter_record(w.name, hops_per_launch)
    traffic = ctr["bytes_per_launch"] if ctr else None
    valu = None
    if ctr:
        valu = {
            "insts_per_wave_hop": ctr["valu_insts_per_wave_hop"],
            "issue_frac": ctr["valu_insts"] * 4 / (1024 * 2.4e9 * avg_ms * 1e-3),
            "l2_requests_per_hop": ctr["l2_requests_per_hop"],
            "wave_wait_frac": (ctr["wait_any"] / ctr["wave_cycles"]
                               if ctr.get("wait_any") and ctr.get("wave_cycles") else None),
            "note": "SQ_INSTS_VALU / (hops / 64); x 4 cycles / (1,024 SIMDs x 2.4 GHz x this "
                    "run's launch time); TCC_REQ_sum / hops; SQ_WAIT_ANY / SQ_WAVE_CYCLES (the "
                    "share of a wave's resident time spent waiting, mostly on its row load)",
        }
    sur_ctr = counter_record("surrogate", None)
    if sur_ctr and sur_roof:
        sur_roof["mfma_flop_counted"] = sur_ctr.get("mfma_flop")
        sur_roof["mfma_busy_cycles"] = sur_ctr.get("mfma_busy_cycles")
        if sur_ctr.get("mfma_flop"):
            sur_roof["mfma_tflops_counted"] = sur_ctr["mfma_flop"] / (sur_roof["avg_launch_ms"]
                                                                      * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(net, w, tables, csr)
        cpu["table_build"] = cpu_table_baseline(tabs)

    value = hops_total / wall
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "hops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": w.name,
            "sites": int(net["n"]),
            "pairs": int(nnz),
            "excitons": E,
            "hops_per_exciton": STEP_HOPS * args.steps,
            "hops_per_step": STEP_HOPS,
            "cutoff_nm": w.cutoff * 1e9,
            "chiralities": [list(c) for c in w.chiralities],
            "rates": ("first-order transfer rates at 300 K, 9 donor/acceptor tables on the "
                      "input.json grid, built by the BF16 surrogate with exact fallback"),
            "msd_bins": args.bins,
        },
        "rate_table_build_s": {
            "mlp": tabs["mlp_s"],
            "exact": tabs["exact_s"],
            "speedup": tabs["exact_s"] / tabs["mlp_s"] if tabs["exact_s"] else None,
            "nodes": tabs["nodes"],
            "nodes_exact_fallback": tabs["fallback"],
            "rel_err_median": tabs["err_median"],
            "rel_err_p99": tabs["err_p99"],
            "frac_over_5pct": tabs["err_over"],
            "note": "exact = first-order rate on the GPU for every node; mlp = fused BF16 "
                    "inference plus exact recomputation of flagged nodes; wall time incl. "
                    "host<->device copies",
        },
        "cnt_model_build_s": dict(tabs["cnt_s"], note=(
            "exciton dispersion of the film's chiralities (calculate_exciton_dispersion): the "
            "Ohno transform v(q) and polarization PI(q) on the GPU vs the whole build on the "
            "host cores; setup, outside the timed region")),
        "csr_build_s": build_s,
        "diffusion_m2_per_s": D,
        "diffusion_ci95": ci,
        "diffusion_exact_tables_m2_per_s": Dx,
        "diffusion_exact_tables_ci95": cix,
        "diffusion_rel_diff_vs_exact": d_rel,
        "reduce_kernel_ms": red_ms,
        "roofline": {
            "kernel": "k_hop<count,sample>",
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "traffic_counters": ("profiles/hop_kernel_counters.json (sources hash "
                                 f"{ctr['source_hash'][:12]})" if ctr else
                                 "none taken on these kernel sources"),
            "valu": valu,
            "bytes_per_hop": bytes_per_hop,
            "visit_weighted_degree": d_v,
            "avg_launch_ms": avg_ms,
            "launches_per_step": launches,
            "hops_per_launch": hops_per_launch,
        },
        "roofline_surrogate": sur_roof,
        "ensemble": ens,
        "contact_mode": contact,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
