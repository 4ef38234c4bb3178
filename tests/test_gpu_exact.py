"""GPU tests of the exact first-order transfer rate (exciton_transfer::first_order on the device)
against the oracle's literal CPU restatement (per-pair double-loop J), against the far-field
closed form, and wired into a KMC rate build.  Tolerance: 1e-9 relative (the device sums the
4.75 M Coulomb terms of a (7,5) pair in a different order than the reference's double loop)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
NM = 1e-9
GRID_LO = (1.5e-9, -10e-9, -10e-9)   # input.json:161-163 table box: z, axis shift 1, axis shift 2
GRID_HI = (10e-9, 10e-9, 10e-9)


def say(*a):
    print("[gpu-exact]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def models():
    from cnt_film_monte_carlo_amd import exciton as X
    return {c: X.CntModel(*c, 10) for c in [(4, 2), (7, 5), (8, 6), (9, 7)]}


@pytest.mark.parametrize("pair,T,geoms", [
    (((4, 2), (4, 2)), 300.0, [(1.5e-9, 0, 0, 0.0), (3e-9, 2e-9, -1e-9, 1.2), (2e-9, -5e-9, 4e-9, 3.141592),
                               (7e-9, 0, 0, 1.5707963), (1.6e-9, 1e-9, 1e-9, 2.3561944)]),
    (((7, 5), (7, 5)), 300.0, [(1.5e-9, 0, 0, 0.0), (4e-9, -3e-9, 2e-9, 0.7)]),
    (((8, 6), (7, 5)), 300.0, [(2.5e-9, 1e-9, 0, 2.0)]),
    # 19 acceptor momenta: two kernel passes (16 + 3)
    (((7, 5), (9, 7)), 400.0, [(2e-9, 1e-9, -2e-9, 0.9)]),
])
def test_transfer_matches_literal_first_order(models, orc, pair, T, geoms):
    from cnt_film_monte_carlo_amd import exciton as X
    d, a = pair
    say(f"first_order {d}->{a} T={T} at {len(geoms)} geometries")
    t = X.Transfer(models[d], models[a], T)
    g = np.array(geoms)
    got = t.rates(g[:, 3], g[:, 0], g[:, 1], g[:, 2])
    td, ta = orc.tube_from_model(models[d]), orc.tube_from_model(models[a])
    for k, (z, s1, s2, th) in enumerate(geoms):
        ref, _ = orc.first_order(td, ta, T, X.BROADENING, z, s1, s2, th)
        assert got[k] == pytest.approx(ref, rel=1e-9, abs=1e-300), (k, got[k], ref)


@pytest.mark.parametrize("pair,T", [(((7, 5), (8, 6)), 300.0), (((7, 5), (9, 7)), 400.0)])
def test_split_launch_carries_the_same_bits(models, pair, T):
    # fewer than 4096 geometries: one workgroup per (geometry, 256-atom donor chunk) and a
    # finishing kernel; 4096 or more: one workgroup per geometry.  Same sums in the same order,
    # so the same bits (the second pair runs two passes of acceptor momenta: 16 + 3).
    from cnt_film_monte_carlo_amd import exciton as X
    t = X.Transfer(models[pair[0]], models[pair[1]], T)
    rng = np.random.default_rng(3)
    n = 4200
    th = rng.uniform(0, np.pi, n)
    z = rng.uniform(1.5e-9, 10e-9, n)
    a1 = rng.uniform(-10e-9, 10e-9, n)
    a2 = rng.uniform(-10e-9, 10e-9, n)
    say(f"split vs whole launches {pair} T={T}: {n} geometries")
    whole = t.rates(th, z, a1, a2)
    idx = rng.choice(n, 37, replace=False)
    part = t.rates(th[idx], z[idx], a1[idx], a2[idx])
    assert np.all(whole > 0)
    assert np.array_equal(part, whole[idx])


def test_no_matched_pairs_gives_zero(models):
    # (9,7) donor to (7,5) acceptor at 150 K: no state pair within the Lorentzian window
    from cnt_film_monte_carlo_amd import exciton as X
    t = X.Transfer(models[(9, 7)], models[(7, 5)], 150.0)
    assert t.info.n_pairs == 0 and t.info.n_terms == 0
    r = t.rates([0.3, 1.0], [2e-9, 5e-9], [0.0, 1e-9], [0.0, -1e-9])
    assert np.all(r == 0.0)


def test_far_field_is_dipolar(models):
    # Zero-charge transition densities: J ~ cos(theta) / z^3 far from the tubes when the
    # separation is perpendicular to both axes, so rate * z^6 -> const, rate ~ cos^2(theta).
    from cnt_film_monte_carlo_amd import exciton as X
    m = models[(7, 5)]
    t = X.Transfer(m, m, 300.0)
    th = np.array([0.0, 0.0, 0.6, 1.2, 2.5, math.pi / 2])
    z = np.array([2e-6] + [4e-6] * 5)
    r = t.rates(th, z, np.zeros(6), np.zeros(6))
    assert r[0] / r[1] == pytest.approx(64.0, rel=1e-3)
    np.testing.assert_allclose(r[2:5] / r[1], np.cos(th[2:5]) ** 2, rtol=1e-5)
    assert r[5] / r[1] < 1e-9


def test_exact_rates_build_and_hop_parity(kmc, orc, nw, models):
    from cnt_film_monte_carlo_amd import exciton as X
    w = nw.film_small()
    net = nw.build_network(w)
    ctx = nw.make_context(w, net, seed=31)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_PAIR)
    geo = ctx.get_pair_geometry()
    t = X.Transfer(models[(7, 5)], models[(7, 5)], 300.0)
    t.set_clamp(GRID_LO, GRID_HI)
    ms = ctx.build_rates_exact([t])
    say(f"exact rates for {ctx.nnz} pairs of the 5k-site film: {ms:.1f} ms")
    rates = ctx.get_pair_rates()
    assert np.all(np.isfinite(rates)) and np.all(rates >= 0)
    # spot-check pairs against the literal CPU first_order at the clamped geometry
    td = orc.tube_from_model(models[(7, 5)])
    rng = np.random.default_rng(2)
    for p in rng.choice(ctx.nnz, 6, replace=False):
        z = min(max(geo["z"][p], GRID_LO[0]), GRID_HI[0])
        s1 = min(max(geo["a1"][p], GRID_LO[1]), GRID_HI[1])
        s2 = min(max(geo["a2"][p], GRID_LO[2]), GRID_HI[2])
        ref, _ = orc.first_order(td, td, 300.0, X.BROADENING, z, s1, s2, geo["theta"][p])
        assert rates[p] == pytest.approx(ref, rel=1e-9), p
    # the CSR holds their FP64 prefix sums in reference order; hops stay bit-exact
    g = ctx.get_csr()
    onet = orc.Net(net, w.cutoff, 0.4 * NM, kmc.PERIODIC_Y, (0, 0, 0), (0, w.period_y[1], 0))
    o = onet.csr(mode=1, pair_rates=rates)
    assert np.array_equal(g["cum"], o["cum"]) and np.array_equal(g["col"], o["col"])
    E, H = 512, 4000
    ctx.inject(E)
    ctx.run(H)
    ox = orc.Excitons(g, net["flags"], seed=31, n_excitons=E)
    ox.run(H)
    st = ctx.get_excitons()
    assert np.array_equal(st["site"], ox.site) and np.array_equal(st["hash"], ox.hash)
    fin = np.isfinite(ox.t_next)
    np.testing.assert_allclose(st["t_next"][fin], ox.t_next[fin], rtol=1e-12)
    ctx.close()
