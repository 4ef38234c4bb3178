"""GPU parity of the HIP path (libcntkmc.so, through the C ABI) against the CPU oracle on
the same seeded inputs.  Bar: integer/index results bit-exact; FP64 cumulative rates and
displacements bit-exact (contraction off on both sides); times within 1e-12 relative (device
log() may differ from glibc by an ulp); reductions within 1e-12 relative."""
import math
import sys

import numpy as np
import pytest

from conftest import site_arrays

pytestmark = pytest.mark.gpu
NM = 1e-9
T_REL = 1e-12


def say(*a):
    print("[gpu]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def small(kmc, orc, nw):
    w = nw.film_small()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(1)
    onet = orc.Net(net, w.cutoff, 0.4 * NM, kmc.PERIODIC_Y, (0, 0, 0), (0, w.period_y[1], 0))
    onet.set_table(th, z, a1, a2, rates)
    ocsr = onet.csr(0)
    return dict(w=w, net=net, tables=(th, z, a1, a2, rates), onet=onet, ocsr=ocsr)


def _ctx(kmc, nw, small, seed=7):
    ctx = nw.make_context(small["w"], small["net"], seed=seed)
    th, z, a1, a2, rates = small["tables"]
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    return ctx


def test_cells_match_oracle(kmc, nw, small):
    say("cells: film_5k_sites")
    ctx = nw.make_context(small["w"], small["net"])
    ctx.build_cells()
    dims, start, sites = ctx.get_cells()
    odims, ostart, osites = small["onet"].cells()
    assert np.array_equal(dims, odims)
    assert np.array_equal(start, ostart)
    assert np.array_equal(sites, osites)
    ctx.close()


def test_csr_table_bitwise(kmc, nw, small):
    say("csr table: film_5k_sites")
    ctx = _ctx(kmc, nw, small)
    g = ctx.get_csr()
    o = small["ocsr"]
    assert np.array_equal(g["row_ptr"], o["row_ptr"])
    assert np.array_equal(g["col"], o["col"])
    assert np.array_equal(g["cum"], o["cum"])                 # bitwise FP64
    assert np.array_equal(g["disp"], o["disp"])
    last = g["row_ptr"][1:] - 1
    has = np.diff(g["row_ptr"]) > 0
    assert np.array_equal(g["gamma"][has], g["cum"][last[has]])
    ctx.close()


def test_pair_geometry_and_pair_rates_bitwise(kmc, orc, nw, small):
    say("pair geometry: film_5k_sites")
    ctx = nw.make_context(small["w"], small["net"])
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_PAIR)
    geo = ctx.get_pair_geometry()
    og = small["onet"].csr(mode=1, pair_rates=None, geometry=True)
    for k in ("z", "a1", "a2"):
        assert np.array_equal(geo[k], og[k]), k
    # theta = acos(cos theta): the device acos and glibc's may differ by one ulp
    th_g, th_o = geo["theta"], og["theta"]
    ulp = np.abs(th_g.view(np.int64) - th_o.view(np.int64))
    say(f"  theta: {np.count_nonzero(ulp)} of {ulp.size} differ, max {ulp.max()} ulp")
    assert ulp.max() <= 1
    rng = np.random.default_rng(4)
    rates = rng.random(ctx.nnz) * 1e12
    rates[::7] = 0.0
    ctx.set_pair_rates(rates)
    g = ctx.get_csr()
    o = small["onet"].csr(mode=1, pair_rates=rates)
    assert np.array_equal(g["cum"], o["cum"])
    ctx.close()


def _compare_state(g, o, label):
    assert np.array_equal(g["site"], o.site), label
    assert np.array_equal(g["hops"], o.hops), label
    assert np.array_equal(g["hash"], o.hash), label
    assert np.array_equal(g["drains"], o.drains), label
    assert np.array_equal(g["disp"], o.dsp), label
    fin = np.isfinite(o.t_next)
    assert np.array_equal(np.isfinite(g["t_next"]), fin)
    np.testing.assert_allclose(g["t_next"][fin], o.t_next[fin], rtol=T_REL, atol=0)


def test_hops_bitwise_small_film(kmc, orc, nw, small):
    say("hops: film_5k_sites 1000 x 10000")
    ctx = _ctx(kmc, nw, small, seed=7)
    E, H = 1000, 10_000
    ctx.inject(E)
    csr = ctx.get_csr()
    ox = orc.Excitons(csr, small["net"]["flags"], seed=7, n_excitons=E)
    g0 = ctx.get_excitons()
    _compare_state(g0, ox, "after injection")
    t_s = float(np.median(ox.t_next)) * 2000
    ctx.set_samples(32, t_s)
    ox.set_samples(32, t_s)
    ms = ctx.run(H)
    ox.run(H)
    say(f"  kernel {ms:.3f} ms")
    _compare_state(ctx.get_excitons(), ox, "after run")
    assert np.array_equal(ctx.get_hop_counts(), ox.site_count)
    assert int(ox.site_count.sum()) == E * H
    s = ctx.get_samples()
    assert np.array_equal(s["n_recorded"], ox.k_rec)
    valid = np.arange(32)[:, None] < ox.k_rec[None, :]
    say(f"  MSD records: {valid.sum()} of {valid.size} reached")
    assert np.array_equal(s["disp"][valid], ox.rec_disp[valid])
    assert np.array_equal(s["drains"][valid], ox.rec_drain[valid])
    assert np.all(s["disp"][~valid] == 0) and np.all(s["drains"][~valid] == 0)
    red, _ = ctx.reduce_samples(16)
    ored = ox.reduce_samples(16)
    np.testing.assert_allclose(red, ored, rtol=1e-12, atol=0)
    ctx.close()


def test_resume_and_trace(kmc, orc, nw, small):
    say("resume + trace: film_5k_sites 256 x 3000")
    ctx = _ctx(kmc, nw, small, seed=11)
    E = 256
    ctx.inject(E)
    ctx.set_trace(8, 4000)
    for h in (1000, 1500, 500):
        ctx.run(h)
    csr = ctx.get_csr()
    ox = orc.Excitons(csr, small["net"]["flags"], seed=11, n_excitons=E)
    ox.set_trace(8, 4000)
    ox.run(3000)
    _compare_state(ctx.get_excitons(), ox, "resume")
    tr = ctx.get_trace()
    assert np.array_equal(tr["n_hops"], ox.tr_n)
    assert np.array_equal(tr["sites"], ox.tr_site)
    np.testing.assert_allclose(tr["times"], ox.tr_time, rtol=T_REL, atol=0)
    ctx.close()


def test_time_cap(kmc, orc, nw, small):
    say("t_max cap: film_5k_sites")
    ctx = _ctx(kmc, nw, small, seed=5)
    E = 512
    ctx.inject(E)
    csr = ctx.get_csr()
    ox = orc.Excitons(csr, small["net"]["flags"], seed=5, n_excitons=E)
    t_cap = float(np.median(ox.t_next)) * 300
    ctx.run(100_000, t_max=t_cap)
    ox.run(100_000, t_max=t_cap)
    g = ctx.get_excitons()
    _compare_state(g, ox, "t_max")
    assert np.all(g["t_next"] > t_cap)
    ctx.close()


def test_time_cap_in_calls_with_trace_and_samples(kmc, orc, nw, small):
    """A clock-bounded run in three calls of rising t_max with the trace and MSD samples on.
    Each launch leaves out the excitons whose next hop falls after t_max and that owe no record
    before it (they sort last and the hop kernel deals only the others); state, per-site counts,
    trace and records equal the oracle's."""
    say("t_max in three calls with trace and samples: film_5k_sites 512 excitons")
    ctx = _ctx(kmc, nw, small, seed=17)
    E = 512
    ctx.inject(E)
    csr = ctx.get_csr()
    ox = orc.Excitons(csr, small["net"]["flags"], seed=17, n_excitons=E)
    T = float(np.median(ox.t_next)) * 300
    for o in (ctx, ox):
        o.set_samples(24, T)
        o.set_trace(8, 50_000)
    for f in (0.25, 0.6, 1.0):
        ctx.run(1_000_000, t_max=f * T)
        ox.run(1_000_000, t_max=f * T)
        _compare_state(ctx.get_excitons(), ox, f"t_max {f} T")
    assert np.array_equal(ctx.get_hop_counts(), ox.site_count)
    tr = ctx.get_trace()
    assert np.array_equal(tr["n_hops"], ox.tr_n)
    assert np.array_equal(tr["sites"], ox.tr_site)
    np.testing.assert_allclose(tr["times"], ox.tr_time, rtol=T_REL, atol=0)
    s = ctx.get_samples()
    assert np.array_equal(s["n_recorded"], ox.k_rec)
    valid = np.arange(24)[:, None] < ox.k_rec[None, :]
    assert np.array_equal(s["disp"][valid], ox.rec_disp[valid])
    assert np.array_equal(s["drains"][valid], ox.rec_drain[valid])
    g = ctx.get_excitons()
    say(f"  {int(g['hops'].sum())} hops, {int(valid.sum())} records")
    ctx.close()


def test_large_film_time_cap(kmc, orc, nw):
    """The 625k-site film, 10^4 excitons run to t_max = 300 median dwell times in two calls with
    MSD samples on (up to ~4e4 hops for one exciton, so most launches leave most excitons out as
    idle and deal the rest over the XCDs): state, per-site counts and records equal the oracle's."""
    say("large film, clock-bounded: 10^4 excitons to 300 median dwell times")
    w = nw.film_large()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(3)
    ctx = nw.make_context(w, net, seed=23)
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    E = 10_000
    ctx.inject(E)
    ox = orc.Excitons(ctx.get_csr(), net["flags"], seed=23, n_excitons=E)
    T = float(np.median(ox.t_next)) * 300
    ctx.set_samples(16, T)
    ox.set_samples(16, T)
    for f in (0.5, 1.0):
        ctx.run(10_000_000, t_max=f * T)
        ox.run(10_000_000, t_max=f * T)
        _compare_state(ctx.get_excitons(), ox, f"large film, t_max {f} T")
    assert np.array_equal(ctx.get_hop_counts(), ox.site_count)
    s = ctx.get_samples()
    assert np.array_equal(s["n_recorded"], ox.k_rec)
    valid = np.arange(16)[:, None] < ox.k_rec[None, :]
    assert np.array_equal(s["disp"][valid], ox.rec_disp[valid])
    say(f"  {int(ox.hops.sum())} hops, launches {ctx.run_stats()[0]}")
    ctx.close()


def test_trapped_exciton(kmc, orc):
    say("trapped: isolated source site")
    # site 0 is isolated (no neighbour within the cutoff) and the only source
    pos = np.array([[0, 0, 0], [100 * NM, 0, 0], [105 * NM, 0, 0]], float)
    ori = np.tile([1.0, 0.0, 0.0], (3, 1))
    net = site_arrays(pos, ori, flags=np.array([1, 0, 2], np.uint8))
    ctx = kmc.Context(kmc.Config(cutoff=20 * NM, seed=1))
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"],
                  flags=net["flags"])
    ctx.set_rate_table([0.0], [0.0], [0.0], [0.0], np.full(1, 1e12))
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    ctx.inject(64)
    ctx.set_samples(4, 1e-9)
    ctx.run(100)
    g = ctx.get_excitons()
    assert np.all(g["site"] == 0) and np.all(g["hops"] == 0)
    assert np.all(np.isinf(g["t_next"]))
    s = ctx.get_samples()
    assert np.all(s["n_recorded"] == 4) and np.all(s["disp"] == 0)
    ctx.close()


def test_ring_known_diffusion_gpu(kmc):
    say("ring: D = k a^2")
    n, a, k = 8000, 10 * NM, 1e12
    pos = np.zeros((n, 3))
    pos[:, 1] = np.arange(n) * a
    ori = np.tile([0.0, 1.0, 0.0], (n, 1))
    ctx = kmc.Context(kmc.Config(cutoff=1.5 * a, periodic_mask=kmc.PERIODIC_Y,
                                 period_len=(0, n * a, 0), seed=3))
    ctx.set_sites(pos[:, 0], pos[:, 1], pos[:, 2], ori[:, 0], ori[:, 1], ori[:, 2],
                  flags=np.ones(n, np.uint8))
    ctx.set_rate_table([0.0], [0.0], [0.0], [0.0], np.full(1, k))
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    E = 20000
    ctx.inject(E)
    t_s = 4e-10
    ctx.set_samples(20, t_s)
    ctx.run(5000)
    red, _ = ctx.reduce_samples(32)
    assert np.all(red[:, :, 0].sum(1) == E)
    D, ci = kmc.estimate_diffusion(red, t_s)
    say(f"  D_y = {D[1]:.4e} +- {ci[1]:.2e}  (exact {k * a * a:.4e})")
    assert abs(D[1] - k * a * a) < 3 * ci[1]
    assert D[0] == 0 and D[2] == 0
    ctx.close()


def test_error_statuses(kmc, nw, small):
    say("error statuses")
    ctx = nw.make_context(small["w"], small["net"])
    with pytest.raises(kmc.KmcError) as e:
        ctx.run(10)
    assert e.value.code == kmc.ERR_STATE
    with pytest.raises(kmc.KmcError) as e:
        ctx.set_rate_table([0.0], [0.0], [0.0], [0.0], np.full(1, -1.0))
    assert e.value.code == kmc.ERR_RATE
    with pytest.raises(kmc.KmcError) as e:
        ctx.build_rates(kmc.RATE_TABLE)
    assert e.value.code == kmc.ERR_STATE
    net = small["net"]
    with pytest.raises(kmc.KmcError) as e:
        ctx.set_sites(net["x"], net["y"] + 600 * NM, net["z"], net["ox"], net["oy"], net["oz"])
    assert e.value.code == kmc.ERR_GEOMETRY
    ctx.close()


def test_large_film_csr_and_hops(kmc, orc, nw):
    say("large film: ~625k sites, 3 chiralities, cutoff 20 nm")
    w = nw.film_large()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(3)
    # make the three chiralities distinguishable: scale table (ci, cj) by (1 + ci + 2 cj)
    scale = np.array([1 + ci + 2 * cj for ci in range(3) for cj in range(3)], float)
    rates = rates * scale[:, None, None, None, None]
    ctx = nw.make_context(w, net, seed=21)
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    g = ctx.get_csr()
    onet = orc.Net(net, w.cutoff, 0.4 * NM, kmc.PERIODIC_Y, (0, 0, 0), (0, w.period_y[1], 0))
    onet.set_table(th, z, a1, a2, rates)
    o = onet.csr(0)
    say(f"  sites {net['n']}, pairs {g['col'].size}")
    assert np.array_equal(g["row_ptr"], o["row_ptr"])
    assert np.array_equal(g["col"], o["col"])
    assert np.array_equal(g["cum"], o["cum"])
    assert np.array_equal(g["disp"], o["disp"])
    # the 10^4-exciton subset on which the full-size film is held bit-exact
    E, H = 10_000, 5000
    ctx.inject(E)
    ctx.run(H)
    ox = orc.Excitons(g, net["flags"], seed=21, n_excitons=E)
    ox.run(H)
    _compare_state(ctx.get_excitons(), ox, "large film")
    assert np.array_equal(ctx.get_hop_counts(), ox.site_count)
    # the full exciton count (10^6): trajectories are keyed by exciton id, so the first 10^4
    # of the full run are the subset's, bit for bit; the run split in two (with the exciton
    # sorts between sub-launches) equals one call; drains + in flight account for every exciton
    say("  10^6 excitons x 5000 hops: first 10^4 vs the oracle, split vs whole")
    Ef = 1_000_000
    ctx.set_hop_schedule(625)
    ctx.inject(Ef)
    ctx.run(H)
    whole = ctx.get_excitons()
    counts_whole = ctx.get_hop_counts()
    sub = {k: v[:E] for k, v in whole.items()}
    _compare_state(sub, ox, "large film, first 10^4 of 10^6")
    ctx.inject(Ef)
    ctx.run(H // 2)
    ctx.run(H - H // 2)
    split = ctx.get_excitons()
    for k in ("site", "hops", "hash", "drains", "disp"):
        assert np.array_equal(split[k], whole[k]), k
    assert np.array_equal(split["t_next"], whole["t_next"])
    # every hop counted once at its source site (the counts restart at each injection)
    assert int(counts_whole.sum()) == int(whole["hops"].sum()) > 0.99 * Ef * H
    ctx.close()
