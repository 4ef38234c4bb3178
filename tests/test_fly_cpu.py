"""Pins of the oracle's drift along the tubes (particle::fly, src/monte_carlo/particle.cpp:9-54)
for v > 0, against closed forms: free flight on a straight tube reflects at the tube ends
(a triangle wave in time), a drain at the far end turns it into uniform motion (displacement
v t, floor(v t / L) drains, monte_carlo.cpp:330-336 keeps the displacement), an isolated site
does not move (particle.cpp:11-12), a periodic axis is crossed by its minimum image; plus
v -> 0 reproducing the pure hopping run and split runs equal to one run."""
import math

import numpy as np
import pytest

NM = 1e-9


def chain_csr(pos, period=(0.0, 0.0, 0.0)):
    """A tube of sites with no hopping neighbours (every dwell infinite)."""
    n = len(pos)
    return dict(row_ptr=np.zeros(n + 1, np.int64), col=np.zeros(0, np.int32), cum=np.zeros(0),
                site_pos=np.asarray(pos, np.float32), period=np.asarray(period, float))


def chain_links(n):
    left = np.arange(-1, n - 1, dtype=np.int32)
    right = np.arange(1, n + 1, dtype=np.int32)
    right[-1] = -1
    return left, right


def straight_tube(n, a, axis=0, start=0.0, period=0.0):
    pos = np.zeros((n, 3))
    pos[:, axis] = start + a * np.arange(n)
    if period > 0:
        pos[:, axis] = np.mod(pos[:, axis], period)
    per = [0.0, 0.0, 0.0]
    per[axis] = period
    return pos, per


def free_flight(orc, pos, per, flags, v, n_bins, t_s, E=4, seed=11):
    ex = orc.Excitons(chain_csr(pos, per), flags, seed, E, velocity=v,
                      links=chain_links(len(pos)))
    assert np.all(np.isinf(ex.t_next))
    ex.set_samples(n_bins, t_s)
    ex.run(100)
    assert np.all(ex.k_rec == n_bins)
    tk = (np.arange(n_bins) + 1) * (t_s / n_bins)
    return ex, tk


def triangle(u, L):
    m = np.mod(u, 2 * L)
    return np.where(m <= L, m, 2 * L - m)


def test_free_flight_reflects_at_tube_ends(orc):
    n, a, v = 11, 10 * NM, 1.0e3
    pos, per = straight_tube(n, a)
    f32 = pos.astype(np.float32).astype(np.float64)
    L = f32[-1, 0] - f32[0, 0]
    flags = np.zeros(n, np.uint8)
    flags[0] = 1
    t_s = 7.3 * L / v
    ex, tk = free_flight(orc, pos, per, flags, v, 97, t_s)
    want = triangle(v * tk, L)
    for e in range(ex.E):
        np.testing.assert_allclose(ex.rec_disp[:, e, 0], want, rtol=0, atol=1e-12 * v * t_s)
        assert np.all(ex.rec_disp[:, e, 1:] == 0.0)
    assert np.all(ex.rec_drain == 0)


def test_drain_at_tube_end_gives_uniform_motion(orc):
    n, a, v = 9, 10 * NM, 2.0e2
    pos, per = straight_tube(n, a)
    f32 = pos.astype(np.float32).astype(np.float64)
    L = f32[-1, 0] - f32[0, 0]
    flags = np.zeros(n, np.uint8)
    flags[0] = 1
    flags[-1] = 2
    t_s = 5.55 * L / v
    ex, tk = free_flight(orc, pos, per, flags, v, 50, t_s)
    u = v * tk
    frac = np.mod(u / L, 1.0)
    keep = (frac > 1e-6) & (frac < 1 - 1e-6)          # away from the drain instants
    for e in range(ex.E):
        np.testing.assert_allclose(ex.rec_disp[keep, e, 0], u[keep], rtol=1e-12, atol=0)
        np.testing.assert_array_equal(ex.rec_drain[keep, e], np.floor(u[keep] / L))


def test_periodic_axis_crossed_by_minimum_image(orc):
    P = 100 * NM
    n, a, v = 6, 10 * NM, 5.0e2
    pos, per = straight_tube(n, a, axis=1, start=75 * NM, period=P)
    assert pos[3, 1] < pos[2, 1]                        # the tube wraps between sites 2 and 3
    f32 = pos.astype(np.float32).astype(np.float64)
    L = sum(orc_min(f32[k + 1, 1] - f32[k, 1], P) for k in range(n - 1))
    flags = np.zeros(n, np.uint8)
    flags[0] = 1
    t_s = 3.1 * L / v
    ex, tk = free_flight(orc, pos, per, flags, v, 40, t_s)
    for e in range(ex.E):
        np.testing.assert_allclose(ex.rec_disp[:, e, 1], triangle(v * tk, L), rtol=0,
                                   atol=1e-12 * v * t_s)


def orc_min(d, P):
    if d > 0.5 * P:
        return d - P
    if d < -0.5 * P:
        return d + P
    return d


def test_isolated_site_does_not_move(orc):
    pos = np.array([[0.0, 0.0, 0.0], [40 * NM, 0.0, 0.0]])
    csr = chain_csr(pos)
    flags = np.array([1, 0], np.uint8)
    ex = orc.Excitons(csr, flags, 3, 2, velocity=1e4,
                      links=(np.array([-1, -1], np.int32), np.array([-1, -1], np.int32)))
    ex.set_samples(8, 1e-9)
    ex.run(10)
    assert np.all(ex.rec_disp == 0.0) and np.all(ex.k_rec == 8)


def test_pass_bound_raises(orc):
    n, a = 3, 10 * NM
    pos, per = straight_tube(n, a)
    flags = np.zeros(n, np.uint8)
    flags[0] = 1
    ex = orc.Excitons(chain_csr(pos, per), flags, 5, 1, velocity=1.0,
                      links=chain_links(n))
    ex.set_samples(1, 1.0)                              # 1e8 site passes > 2^22
    with pytest.raises(RuntimeError, match="2\\^22"):
        ex.run(1)


# --------------------------------------------------------------------------------------------
# Drift together with hops on the 5k-site film (Förster/Wong table)
# --------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def film(orc, kmc, nw):
    w = nw.film_small()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(1)
    onet = orc.Net(net, w.cutoff, 0.4e-9, kmc.PERIODIC_Y, (0, 0, 0), (0, w.period_y[1], 0))
    onet.set_table(th, z, a1, a2, rates)
    return net, onet.csr(0)


def _film_excitons(orc, film, v, E=128, seed=21):
    net, csr = film
    return orc.Excitons(csr, net["flags"], seed, E, velocity=v, links=(net["left"], net["right"]))


def test_vanishing_velocity_reproduces_hopping(orc, film):
    H = 1500
    ref = _film_excitons(orc, film, 0.0)
    ref.run(H)
    slow = _film_excitons(orc, film, 1e-12)
    slow.run(H)
    for k in ("site", "hash", "hops", "drains", "site_count", "t_next"):
        np.testing.assert_array_equal(getattr(slow, k), getattr(ref, k), err_msg=k)
    np.testing.assert_array_equal(slow.dsp, ref.dsp)
    assert np.all(slow.s_off < 1e-12 * np.max(slow.t_next) * 10)


@pytest.mark.parametrize("v", [1.0, 10.0])
def test_split_runs_equal_one_run(orc, film, v):
    H = 1200
    one = _film_excitons(orc, film, v)
    one.set_samples(16, 1e-6)
    one.run(H)
    two = _film_excitons(orc, film, v)
    two.set_samples(16, 1e-6)
    two.run(H // 3)
    two.run(H - H // 3)
    for k in ("site", "hash", "hops", "drains", "site_count", "t_next", "dsp", "s_off",
              "t_last", "heading", "k_rec", "rec_disp", "rec_drain"):
        np.testing.assert_array_equal(getattr(two, k), getattr(one, k), err_msg=k)


def test_drift_changes_trajectories(orc, film):
    H = 800
    ref = _film_excitons(orc, film, 0.0)
    ref.run(H)
    fast = _film_excitons(orc, film, 10.0)
    fast.run(H)
    # flights pass sites (hash tags them), so the visited sequences differ
    assert np.mean(fast.hash != ref.hash) > 0.9
    assert np.all(fast.hops == H)
