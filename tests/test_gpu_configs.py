"""The film and ensemble configurations at their own sizes on the GPU, against the CPU oracle
(oracle/kmc_oracle.c) and against the product's own single-network runs.

* 256-network ensemble (film_5k_sites rules, 5k-200k sites log-uniform, excitons = 0.2 x sites,
  T uniform in 150-400 K, per-pair surrogate rates at each network's temperature): one
  kmc_ensemble_run of 1,000 hops; three members bit-equal to plain kmc_run; the first 1,000
  excitons of two members bit-equal to the oracle fed the member's CSR.
* 1.6M-site film (80,000 tubes in 20x20x0.5 um, 16M excitons) on surrogate-built tables:
  2,000 hops; the first 10^4 excitons bit-exact against the oracle; the run split in two calls
  equal to one call; departure counts summing to the hops made; a member of a 64-network
  ensemble equal to its plain run.
* 625k-site film: 10^4 excitons x 10^5 hops (10^9 hops) on the Förster tables and on the
  surrogate-built first-order tables the benchmark times, bit-exact against the oracle.

The reference runs one network per process through kubo_step (src/monte_carlo/monte_carlo.cpp:
319-342); the bar is that of the other parity tests: sites, hop counts, visit hashes, drains,
displacements and per-site counts bitwise, times within 1e-12 relative."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
NM = 1e-9
T_REL = 1e-12


def say(*a):
    print("[gpu-configs]", *a, file=sys.stderr, flush=True)


def _compare_state(g, o, label, n=None):
    n = o.E if n is None else n
    sub = {k: v[:n] for k, v in g.items()}
    assert np.array_equal(sub["site"], o.site[:n]), label
    assert np.array_equal(sub["hops"], o.hops[:n]), label
    assert np.array_equal(sub["hash"], o.hash[:n]), label
    assert np.array_equal(sub["drains"], o.drains[:n]), label
    assert np.array_equal(sub["disp"], o.dsp[:n]), label
    fin = np.isfinite(o.t_next[:n])
    assert np.array_equal(np.isfinite(sub["t_next"]), fin), label
    np.testing.assert_allclose(sub["t_next"][fin], o.t_next[:n][fin], rtol=T_REL, atol=0)


def _same(a, b, label):
    for k in ("site", "hops", "hash", "drains", "disp", "t_next"):
        assert np.array_equal(a[k], b[k]), f"{label}: {k}"


@pytest.fixture(scope="module")
def surrogate():
    from cnt_film_monte_carlo_amd import surrogate as S
    sur = S.RateSurrogate.load()
    yield sur
    sur.close()


def _member(kmc, nw, w, net, sur, model, seed):
    """A context of one ensemble network with per-pair surrogate rates at its temperature."""
    from cnt_film_monte_carlo_amd import exciton as X
    tr = X.Transfer(model, model, w.temperature)
    tr.set_clamp(sur.domain.box_lo, sur.domain.box_hi)
    ctx = nw.make_context(w, net, seed=seed)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_PAIR)
    ctx.build_rates_mlp(sur, [0], w.temperature, [tr])
    tr.close()
    return ctx


def _ensemble_check(kmc, orc, nw, sur, n_networks, hops, n_plain, n_oracle, ens_seed=3003):
    from cnt_film_monte_carlo_amd import exciton as X
    t0 = time.perf_counter()
    ws = nw.ensemble(n_networks, seed=ens_seed)
    model = X.CntModel(7, 5, 10, device=0)
    nets = [nw.build_network(w) for w in ws]
    ctxs = [_member(kmc, nw, w, net, sur, model, 5000 + k) for k, (w, net) in
            enumerate(zip(ws, nets))]
    for c, w in zip(ctxs, ws):
        c.inject(w.n_excitons)
    sizes = np.array([c.n_sites for c in ctxs])
    temps = np.array([w.temperature for w in ws])
    say(f"  {n_networks} networks: {sizes.sum()} sites ({sizes.min()}-{sizes.max()}), "
        f"T {temps.min():.0f}-{temps.max():.0f} K, built in {time.perf_counter() - t0:.1f} s")
    assert sizes.min() >= 4_000 and sizes.max() <= 210_000
    ens = kmc.Ensemble(ctxs)
    ms = ens.run(hops)
    n_exc = sum(w.n_excitons for w in ws)
    say(f"  one ensemble run: {n_exc} excitons x {hops} hops in {ms:.1f} ms")
    ens.close()
    order = np.argsort(sizes)
    plain_k = [int(order[0]), int(order[len(order) // 2]), int(order[-1])][:n_plain]
    for k in plain_k:
        c = _member(kmc, nw, ws[k], nets[k], sur, model, 5000 + k)
        c.inject(ws[k].n_excitons)
        c.run(hops)
        _same(ctxs[k].get_excitons(), c.get_excitons(), f"member {k} vs plain")
        assert np.array_equal(ctxs[k].get_hop_counts(), c.get_hop_counts())
        c.close()
    say(f"  members {plain_k} equal their plain runs")
    for k in [int(order[1]), int(order[-2])][:n_oracle]:
        csr = ctxs[k].get_csr()
        n = min(1000, ws[k].n_excitons)
        ox = orc.Excitons(csr, nets[k]["flags"], seed=5000 + k, n_excitons=n)
        ox.run(hops)
        _compare_state(ctxs[k].get_excitons(), ox, f"member {k} vs oracle")
        say(f"  member {k} ({sizes[k]} sites, {ws[k].temperature:.0f} K): first {n} excitons "
            f"equal the oracle")
    for c in ctxs:
        c.close()
    model.close()


def test_ensemble_256_networks(kmc, orc, nw, surrogate):
    say("ensemble_256_networks: 1000 hops in one kmc_ensemble_run")
    _ensemble_check(kmc, orc, nw, surrogate, 256, 1000, 3, 2)


def test_film_1600k_sites_16m_excitons(kmc, orc, nw, surrogate):
    from cnt_film_monte_carlo_amd import surrogate as S
    say("film_1600k_sites: surrogate tables, 16M excitons x 2000 hops")
    t0 = time.perf_counter()
    w = nw.film_xlarge()
    net = nw.build_network(w)
    th, z, a1, a2, rates, n_exact = S.film_tables(w.chiralities, w.temperature, surrogate)
    ctx = nw.make_context(w, net, seed=33)
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    say(f"  {net['n']} sites, {ctx.nnz} pairs, tables ({n_exact} nodes exact) in "
        f"{time.perf_counter() - t0:.1f} s")
    assert net["n"] == 1_600_000
    E, H = w.n_excitons, 2000
    ctx.inject(E)
    ms = ctx.run(H)
    say(f"  whole: {E * H:.2e} hops in {ms:.1f} ms")
    whole = ctx.get_excitons()
    counts = ctx.get_hop_counts()
    assert int(counts.sum()) == int(whole["hops"].sum()) > 0.99 * E * H
    csr = ctx.get_csr()
    n = 10_000
    ox = orc.Excitons(csr, net["flags"], seed=33, n_excitons=n)
    ox.run(H)
    _compare_state(whole, ox, "film_1600k_sites, first 10^4 of 16M")
    say("  first 10^4 excitons equal the oracle")
    ctx.inject(E)
    ctx.run(H // 2)
    ctx.run(H - H // 2)
    _same(ctx.get_excitons(), whole, "split vs whole")
    assert np.array_equal(ctx.get_hop_counts(), counts)
    say("  split run equals one call")
    ctx.close()
    say("film_1600k_sites: the 64-network ensemble")
    _ensemble_check(kmc, orc, nw, surrogate, 64, 1000, 1, 1, ens_seed=3005)


@pytest.mark.parametrize("tables", ["forster", "surrogate"])
def test_film_625k_sites_1e9_hops(kmc, orc, nw, surrogate, tables):
    """10^4 excitons x 10^5 hops on the 625k-site film, bit-exact against the oracle, with MSD
    records and per-site counts."""
    from cnt_film_monte_carlo_amd import surrogate as S
    say(f"film_625k_sites: {tables} tables, 10^4 x 10^5 hops")
    w = nw.film_large()
    net = nw.build_network(w)
    if tables == "forster":
        th, z, a1, a2, rates = nw.forster_tables(3)
        scale = np.array([1 + ci + 2 * cj for ci in range(3) for cj in range(3)], float)
        rates = rates * scale[:, None, None, None, None]
    else:
        th, z, a1, a2, rates, _ = S.film_tables(w.chiralities, w.temperature, surrogate)
    ctx = nw.make_context(w, net, seed=44)
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    csr = ctx.get_csr()
    E, H = 10_000, 100_000
    ctx.inject(E)
    ox = orc.Excitons(csr, net["flags"], seed=44, n_excitons=E)
    t_s = float(np.median(ox.t_next)) * 20_000
    ctx.set_samples(32, t_s)
    ox.set_samples(32, t_s)
    ms = ctx.run(H)
    t0 = time.perf_counter()
    ox.run(H)
    say(f"  GPU {ms:.0f} ms, oracle {time.perf_counter() - t0:.1f} s")
    _compare_state(ctx.get_excitons(), ox, f"film_625k_sites {tables}")
    assert np.array_equal(ctx.get_hop_counts(), ox.site_count)
    assert int(ox.site_count.sum()) == int(ox.hops.sum())
    s = ctx.get_samples()
    assert np.array_equal(s["n_recorded"], ox.k_rec)
    valid = np.arange(32)[:, None] < ox.k_rec[None, :]
    assert valid.sum() > 0
    assert np.array_equal(s["disp"][valid], ox.rec_disp[valid])
    assert np.array_equal(s["drains"][valid], ox.rec_drain[valid])
    say(f"  drains {int(ox.drains.sum())}, MSD records {int(valid.sum())}")
    ctx.close()
