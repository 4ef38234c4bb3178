"""The compiled C++ consumer of include/cntkmc.h (host/kubo_main, built by build() with hipcc
over the monte_carlo host layer) run as a fresh child process: the reference driver's Green-Kubo
loop (src/main.cpp:64-76: kubo_init, kubo_create_particles, then kubo_step(dt) and
kubo_save_avg_dispalcement_squared until kubo_max_time) on the film_5k_sites rules with the Kubo
regions of 5 sections and no periodic axis.  Its final state, per-site counts and every line of
<dx^2>, <dy^2>, <dz^2> are compared with the oracle run step by step to the same clock."""
import math
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "cnt_film_monte_carlo_amd" / "host" / "kubo_main"
NM = 1e-9
BIG = (1 << 63) // 2


def say(*a):
    print("[gpu-host-driver]", *a, file=sys.stderr, flush=True)


def _inputs(kmc, orc):
    p = kmc.NetworkParams(n_tubes=250, len_min=200 * NM, len_max=200 * NM,
                          box_lo=(0.0, 0.0, 0.0), box_hi=(500 * NM, 500 * NM, 50 * NM),
                          site_spacing=10 * NM, n_chir=1, wrap_mask=0, seed=1001)
    net = kmc.generate_network(p)
    net["flags"], _, _ = kmc.kubo_regions(net, 5)
    th, z, a1, a2 = kmc.default_grids()
    rates = kmc.forster_table(1e13, th, z, a1, a2)
    onet = orc.Net(net, 15 * NM, 0.4 * NM, 0)
    onet.set_table(th, z, a1, a2, rates)
    return net, onet.csr(0)


def _read_state(path):
    raw = path.read_bytes()
    E, N = np.frombuffer(raw[:16], np.int64)
    off = 16
    out = {}
    for name, dt, n in (("site", np.int32, E), ("t_next", np.float64, E), ("hops", np.uint64, E),
                        ("disp", np.float64, 3 * E), ("drains", np.uint32, E),
                        ("hash", np.uint64, E), ("counts", np.uint64, N)):
        a = np.frombuffer(raw, dt, int(n), off)
        off += a.nbytes
        out[name] = a
    out["disp"] = out["disp"].reshape(-1, 3)
    assert off == len(raw)
    return out


def test_kubo_main_matches_oracle(kmc, orc, tmp_path):
    assert EXE.exists(), "host/kubo_main is built by build()"
    net, csr = _inputs(kmc, orc)
    E, steps, seed = 2000, 20, 7
    ox = orc.Excitons(csr, net["flags"], seed=seed, n_excitons=E)
    max_time = float(np.median(ox.t_next)) * 300
    say(f"kubo_main: {net['n']} sites, {E} particles, {steps} steps to {max_time:.3e} s")
    res = subprocess.run([str(EXE), str(tmp_path), str(E), str(steps), repr(max_time), str(seed)],
                         capture_output=True, text=True, timeout=300)
    say(res.stdout.strip(), res.stderr.strip())
    assert res.returncode == 0
    g = _read_state(tmp_path / "state.bin")
    # the oracle through the same loop: kubo_step(dt) = every hop due by time() + dt
    dt = max_time / steps
    t = 0.0
    lines = (tmp_path / "particle_dispalcement.avg.squared.dat").read_text().splitlines()
    rows = [ln for ln in lines if ln and ln[0] in "+-0123456789"]
    assert len(rows) == steps
    for k in range(steps):
        t = t + dt
        ox.run(BIG, t_max=t)
        avg = [0.0, 0.0, 0.0]
        for e in range(E):
            for a in range(3):
                avg[a] += float(ox.dsp[e, a]) * float(ox.dsp[e, a])
        avg = [v / E for v in avg]
        got = [float(v) for v in rows[k].split(",")]
        assert got[0] == t
        np.testing.assert_allclose(got[1:], avg, rtol=1e-14, atol=0)
    assert np.array_equal(g["site"], ox.site)
    assert np.array_equal(g["hops"], ox.hops)
    assert np.array_equal(g["hash"], ox.hash)
    assert np.array_equal(g["drains"], ox.drains)
    assert np.array_equal(g["disp"], ox.dsp)
    np.testing.assert_allclose(g["t_next"], ox.t_next, rtol=1e-12, atol=0)
    assert np.array_equal(g["counts"], ox.site_count)
    assert int(ox.drains.sum()) > 0 and int(ox.hops.sum()) > 100 * E
    say(f"  {int(ox.hops.sum())} hops, {int(ox.drains.sum())} re-injections, bit-exact")
