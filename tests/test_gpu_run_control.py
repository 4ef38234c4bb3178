"""GPU tests of how a run ends and what it counts: the early end of kmc_run once no exciton can
hop (the reference's clock-bounded Kubo loop, src/main.cpp:64-76, called with a huge hop budget),
64-bit departure counts past 2^32 per device copy, the nearest-node lookup on exact midpoint
ties in the device CSR build (scattering_struct.h:42-49), the hop-kernel timing of a run, and
MSD records taken window by window."""
import math
import sys

import numpy as np
import pytest

from conftest import site_arrays

pytestmark = pytest.mark.gpu
NM = 1e-9
T_REL = 1e-12
BIG = (1 << 63) // 2


def say(*a):
    print("[gpu-run-control]", *a, file=sys.stderr, flush=True)


def _compare_state(g, o, label):
    assert np.array_equal(g["site"], o.site), label
    assert np.array_equal(g["hops"], o.hops), label
    assert np.array_equal(g["hash"], o.hash), label
    assert np.array_equal(g["drains"], o.drains), label
    assert np.array_equal(g["disp"], o.dsp), label
    fin = np.isfinite(o.t_next)
    assert np.array_equal(np.isfinite(g["t_next"]), fin)
    np.testing.assert_allclose(g["t_next"][fin], o.t_next[fin], rtol=T_REL, atol=0)


def test_huge_budget_ends_at_t_max(kmc, orc, nw):
    """kmc_run(ctx, INT64_MAX / 2, t_cap) with the default schedule returns once every exciton
    is past t_cap, bit-equal to the oracle's run to the same clock.  Launches are 625 hops until
    a read-back finds at most 1/64 of the excitons still able to move, 10,000 hops after that,
    so the few long chains of this film's tail take fewer launches than 625-hop ones would."""
    say("huge budget: film_5k_sites 1000 excitons, t_max cap, budget 2^62")
    w = nw.film_small()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(1)
    ctx = nw.make_context(w, net, seed=13)
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    E = 1000
    ctx.inject(E)
    csr = ctx.get_csr()
    ox = orc.Excitons(csr, net["flags"], seed=13, n_excitons=E)
    t_cap = float(np.median(ox.t_next)) * 400
    ctx.run(BIG, t_max=t_cap)
    launches, live = ctx.run_stats()
    g = ctx.get_excitons()
    # live counts read back after launches 1, 2, 4 and every 8th: at most 8 launches beyond the
    # longest exciton's hops at 625 per launch, at least its hops at 10,000 per launch, and
    # fewer than 625-hop launches would need once the tail is long
    longest = int(g["hops"].max())
    need = -(-longest // 625)
    say(f"  {launches} launches ({need} of 625 hops for the longest, {longest} hops), "
        f"{live} excitons live")
    assert live == 0 and -(-longest // 10_000) <= launches <= need + 8
    if need > 64:
        assert launches < need
    ox.run(BIG, t_max=t_cap)
    _compare_state(g, ox, "huge budget")
    assert np.all(g["t_next"] > t_cap)
    assert np.array_equal(ctx.get_hop_counts(), ox.site_count)
    # a second call has nothing left to do: one launch, no hop
    ctx.run(BIG, t_max=t_cap)
    assert ctx.run_stats() == (1, 0)
    assert np.array_equal(ctx.get_excitons()["hops"], g["hops"])
    ctx.close()


def test_hop_kernel_time(kmc, nw):
    """kmc_get_hop_time: refused before a timed run; after one, positive and at most the
    whole pass's time (the sorts between the sub-launches are left out of it)."""
    say("hop time: film_5k_sites 4096 excitons, 2500 hops in 4 sub-launches")
    w = nw.film_small()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(1)
    ctx = nw.make_context(w, net, seed=5)
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    ctx.inject(4096)
    with pytest.raises(kmc.KmcError):
        ctx.hop_time()
    ms = ctx.run(2500, timed=True)
    hop = ctx.hop_time()
    say(f"  pass {ms:.3f} ms, hop kernels {hop:.3f} ms, {ctx.run_stats(live=False)[0]} launches")
    assert ctx.run_stats(live=False)[0] == 4
    assert 0.0 < hop <= ms
    ctx.run(100, timed=False)
    with pytest.raises(kmc.KmcError):
        ctx.hop_time()


def test_sample_windows_equal_whole_grid(kmc, nw):
    """kmc_set_samples_window: film_5k_sites recorded 16 bins at a time over a 64-bin grid (the
    clock run to each window's last sample time, the window reduced, then the next one opened)
    gives the records, reductions and final state of the whole grid recorded at once, bit for
    bit, with a quarter of the record memory.  Opening the next window before every exciton has
    recorded the open one, or after the last, is refused."""
    say("sample windows: film_5k_sites 3000 excitons, 64 bins as 4 windows of 16")
    w = nw.film_small()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(1)

    def fresh():
        ctx = nw.make_context(w, net, seed=21)
        ctx.set_rate_table(th, z, a1, a2, rates)
        ctx.build_cells()
        ctx.build_rates(kmc.RATE_TABLE)
        ctx.inject(3000)
        return ctx

    a = fresh()
    fin = np.isfinite(a.get_excitons()["t_next"])
    T = float(np.median(a.get_excitons()["t_next"][fin])) * 200
    a.set_samples(64, T)
    a.run(BIG, t_max=T)
    red_a, rec_a = a.reduce_samples(8)[0], a.get_samples()
    b = fresh()
    b.set_samples(64, T, window=16)
    with pytest.raises(kmc.KmcError):
        b.next_sample_window()          # nothing recorded yet
    dt = T / 64
    for i in range(4):
        first, n = b.sample_window()
        assert (first, n) == (16 * i, 16)
        b.run(BIG, t_max=(first + n) * dt)
        red_b, rec_b = b.reduce_samples(8)[0], b.get_samples()
        assert np.array_equal(red_b, red_a[first:first + n])
        assert np.array_equal(rec_b["disp"], rec_a["disp"][first:first + n])
        assert np.array_equal(rec_b["drains"], rec_a["drains"][first:first + n])
        assert np.all(rec_b["n_recorded"] == first + n)
        if i < 3:
            b.next_sample_window()
    with pytest.raises(kmc.KmcError):
        b.next_sample_window()          # the last window is open
    ga, gb = a.get_excitons(), b.get_excitons()
    for k in ("site", "hops", "hash", "drains", "disp", "t_next"):
        assert np.array_equal(ga[k], gb[k]), k
    assert np.array_equal(a.get_hop_counts(), b.get_hop_counts())
    say(f"  {int(gb['hops'].sum())} hops, windows equal the whole grid")
    a.close()
    b.close()


def test_ensemble_huge_budget_ends(kmc, nw):
    say("ensemble huge budget: 3 networks")
    ws = nw.ensemble(n_networks=3, seed=17, n_min=2_000, n_max=8_000)
    th, z, a1, a2, rates = nw.forster_tables(1)
    ctxs = []
    for i, w in enumerate(ws):
        c = nw.make_context(w, nw.build_network(w), seed=300 + i)
        c.set_rate_table(th, z, a1, a2, rates)
        c.build_cells()
        c.build_rates(kmc.RATE_TABLE)
        c.inject(w.n_excitons)
        ctxs.append(c)
    t_cap = 2e-11
    ens = kmc.Ensemble(ctxs)
    ens.run(BIG, t_max=t_cap)
    for c in ctxs:
        t = c.get_excitons()["t_next"]
        assert np.all(t > t_cap)
    ens.close()
    for c in ctxs:
        c.close()


def test_counts_beyond_32_bits(kmc):
    """Two sites that are each other's only neighbour: 10^6 excitons x 10^5 hops leave each site
    5e10 times, about 6e9 per XCD copy, past 2^32.  The copies are folded into 64-bit totals
    before they can wrap, so the counts are exact.  Each workgroup's 16-bit LDS window counter
    of either site takes 80,000 departures per 625-hop launch, so the handover of 0x8000 counts
    to the device copy runs many times as well."""
    E, H = 1_000_000, 100_000
    say(f"ping-pong: 2 sites, {E} x {H} hops")
    pos = np.array([[0.0, 0.0, 0.0], [10 * NM, 0.0, 0.0]])
    ori = np.tile([1.0, 0.0, 0.0], (2, 1))
    net = site_arrays(pos, ori, flags=np.array([1, 0], np.uint8))
    ctx = kmc.Context(kmc.Config(cutoff=15 * NM, seed=9))
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"],
                  flags=net["flags"])
    ctx.set_rate_table([0.0], [0.0], [0.0], [0.0], np.full(1, 1e12))
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    ctx.inject(E)
    ms = ctx.run(H)
    say(f"  {ms:.1f} ms ({E * H / (ms * 1e-3):.3e} hops/s)")
    g = ctx.get_excitons()
    assert np.all(g["hops"] == H)
    counts = ctx.get_hop_counts()
    assert counts.tolist() == [E * ((H + 1) // 2), E * (H // 2)]
    assert counts[0] // 8 > (1 << 32)
    # the run continues and the totals keep adding up
    ctx.run(1001)
    assert ctx.get_hop_counts().tolist() == [E * ((H + 1002) // 2), E * ((H + 1001) // 2)]
    ctx.close()


def _tie_grid(x):
    """Two nodes equally far from x in FP64 (x on the exact midpoint)."""
    for k in range(20, 80):
        d = math.ldexp(1.0, -k)
        lo, hi = x - d, x + d
        if lo != x and hi != x and abs(lo - x) == abs(hi - x):
            return np.array([lo, hi])
    raise AssertionError(f"no exact tie around {x}")


def test_device_lookup_takes_first_node_on_ties(kmc, orc):
    """The CSR build's nearest-node lookup on the device at exact midpoint ties on all four axes:
    the first node, as index_min (scattering_struct.h:42-49) and the oracle give.  Any axis
    taking its second node would pick a different rate."""
    say("device lookup: midpoint ties")
    pos = np.array([[0.0, 0.0, 0.0], [10 * NM, 1 * NM, 0.0]])
    ori = np.tile([1.0, 0.0, 0.0], (2, 1))
    net = site_arrays(pos, ori, flags=np.array([1, 0], np.uint8))
    onet = orc.Net(net, 15 * NM)
    geo = onet.csr(mode=1, pair_rates=np.ones(2), geometry=True)
    q = [float(geo[k][0]) for k in ("theta", "z", "a1", "a2")]   # row of site 0
    say(f"  geometry of 0 -> 1: theta {q[0]}, z {q[1]}, a1 {q[2]}, a2 {q[3]}")
    grids = [_tie_grid(v) for v in q]
    rates = (1.0 + np.arange(16, dtype=float)).reshape(2, 2, 2, 2) * 1e12
    ctx = kmc.Context(kmc.Config(cutoff=15 * NM, seed=1))
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"],
                  flags=net["flags"])
    ctx.set_rate_table(*grids, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    g = ctx.get_csr()
    onet.set_table(*grids, rates)
    o = onet.csr(0)
    assert np.array_equal(g["cum"], o["cum"])
    assert g["cum"][0] == rates[0, 0, 0, 0]
    r_host, nodes = kmc.table_rate(*grids, rates, *[[v] for v in q])
    assert nodes.tolist() == [[0, 0, 0, 0]] and r_host[0] == g["cum"][0]
    ctx.close()
