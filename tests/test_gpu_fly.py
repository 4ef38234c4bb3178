"""GPU parity of the drift along the tubes (particle::fly, src/monte_carlo/particle.cpp:9-54,
kmc_set_velocity) against the oracle on the same seeded inputs.  Bar: sites, visited-site
hashes (hops and site passes), hop and drain counts, per-site departure counts and the
displacement of the current site bit-exact; times, the offset along the current segment and
the MSD records (which hold that offset) within 1e-12 relative plus 1e-11 v t_sample absolute
(the flights last differences of event times, and the device log() may differ from glibc's by
an ulp)."""
import sys

import numpy as np
import pytest

from conftest import site_arrays

pytestmark = pytest.mark.gpu
NM = 1e-9
T_REL = 1e-12


def say(*a):
    print("[gpu-fly]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def small(kmc, orc, nw):
    w = nw.film_small()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(1)
    onet = orc.Net(net, w.cutoff, 0.4 * NM, kmc.PERIODIC_Y, (0, 0, 0), (0, w.period_y[1], 0))
    onet.set_table(th, z, a1, a2, rates)
    return dict(w=w, net=net, tables=(th, z, a1, a2, rates), ocsr=onet.csr(0))


def _ctx(kmc, nw, small, v, seed=7):
    ctx = nw.make_context(small["w"], small["net"], seed=seed)
    th, z, a1, a2, rates = small["tables"]
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    ctx.set_velocity(v)
    return ctx


def _oracle(orc, small, ctx, v, seed, E):
    net = small["net"]
    return orc.Excitons(ctx.get_csr(), net["flags"], seed, E, velocity=v,
                        links=(net["left"], net["right"]))


def _check(g, o, v, t_s):
    st = g.get_excitons()
    fl = g.get_flight()
    assert np.array_equal(st["site"], o.site)
    assert np.array_equal(st["hash"], o.hash)
    assert np.array_equal(st["hops"], o.hops)
    assert np.array_equal(st["drains"], o.drains)
    assert np.array_equal(st["disp"], o.dsp)
    assert np.array_equal(fl["heading"], o.heading)
    assert np.array_equal(g.get_hop_counts(), o.site_count)
    np.testing.assert_allclose(st["t_next"], o.t_next, rtol=T_REL, atol=0)
    np.testing.assert_allclose(fl["t_last"], o.t_last, rtol=T_REL, atol=0)
    np.testing.assert_allclose(fl["offset"], o.s_off, rtol=T_REL, atol=1e-11 * v * t_s)
    if t_s > 0:
        smp = g.get_samples()
        assert np.array_equal(smp["n_recorded"], o.k_rec)
        assert np.array_equal(smp["drains"], o.rec_drain)
        np.testing.assert_allclose(smp["disp"], o.rec_disp, rtol=T_REL, atol=1e-11 * v * t_s)


@pytest.mark.parametrize("v", [0.1, 1.0, 10.0])
def test_film_drift_matches_oracle(kmc, orc, nw, small, v):
    E, H, seed, t_s = 256, 2000, 31, 2e-6
    say(f"film_5k_sites v={v} m/s: {E} excitons x {H} hops")
    ctx = _ctx(kmc, nw, small, v, seed)
    ctx.inject(E)
    ctx.set_samples(24, t_s)
    ctx.run(H)
    o = _oracle(orc, small, ctx, v, seed, E)
    o.set_samples(24, t_s)
    o.run(H)
    _check(ctx, o, v, t_s)
    ctx.close()


def test_film_drift_split_runs_and_schedules(kmc, orc, nw, small):
    v, E, H, seed, t_s = 3.0, 512, 1500, 41, 1e-6
    say(f"film_5k_sites v={v}: split runs and sub-launch schedules")
    ref = _ctx(kmc, nw, small, v, seed)
    ref.set_hop_schedule(0)
    ref.inject(E)
    ref.set_samples(16, t_s)
    ref.run(H)
    ctx = _ctx(kmc, nw, small, v, seed)
    ctx.set_hop_schedule(97)
    ctx.inject(E)
    ctx.set_samples(16, t_s)
    ctx.run(H // 3)
    ctx.run(H - H // 3)
    a, b = ref.get_excitons(), ctx.get_excitons()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    fa, fb = ref.get_flight(), ctx.get_flight()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    assert np.array_equal(ref.get_samples()["disp"], ctx.get_samples()["disp"])
    ref.close()
    ctx.close()


def test_vanishing_velocity_equals_hopping(kmc, nw, small):
    E, H, seed = 512, 2000, 51
    say("film_5k_sites: v = 1e-12 m/s against v = 0")
    a = _ctx(kmc, nw, small, 0.0, seed)
    a.inject(E)
    a.run(H)
    b = _ctx(kmc, nw, small, 1e-12, seed)
    b.inject(E)
    b.run(H)
    ga, gb = a.get_excitons(), b.get_excitons()
    for k in ("site", "hash", "hops", "drains", "disp", "t_next"):
        assert np.array_equal(ga[k], gb[k]), k
    assert np.array_equal(a.get_hop_counts(), b.get_hop_counts())
    a.close()
    b.close()


def _tube_context(kmc, n, a, v, flags, seed=3):
    # one straight tube with spacing above the cutoff: no hopping neighbours, infinite dwells
    pos = np.zeros((n, 3))
    pos[:, 0] = a * np.arange(n) + 20 * NM
    ori = np.tile([1.0, 0.0, 0.0], (n, 1))
    left = np.arange(-1, n - 1, dtype=np.int32)
    right = np.arange(1, n + 1, dtype=np.int32)
    right[-1] = -1
    s = site_arrays(pos, ori, flags=flags, left=left, right=right)
    cfg = kmc.Config(cutoff=0.5 * a, min_distance=0.4e-9, seed=seed)
    ctx = kmc.Context(cfg)
    ctx.set_sites(s["x"], s["y"], s["z"], s["ox"], s["oy"], s["oz"], s["left"], s["right"],
                  s["chirality"], s["flags"])
    th, z, a1, a2 = kmc.default_grids()
    ctx.set_rate_table(th, z, a1, a2, kmc.forster_table(1e13, th, z, a1, a2)[None])
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    ctx.set_velocity(v)
    return ctx, pos


def test_free_flight_closed_forms(kmc):
    n, a, v = 11, 10 * NM, 1.0e3
    say("straight tube: reflection and drain closed forms")
    flags = np.zeros(n, np.uint8)
    flags[0] = 1
    ctx, pos = _tube_context(kmc, n, a, v, flags)
    f32 = pos.astype(np.float32).astype(np.float64)
    L = f32[-1, 0] - f32[0, 0]
    t_s, nb = 7.3 * L / v, 97
    ctx.inject(8)
    ctx.set_samples(nb, t_s)
    ctx.run(10)
    tk = (np.arange(nb) + 1) * (t_s / nb)
    m = np.mod(v * tk, 2 * L)
    want = np.where(m <= L, m, 2 * L - m)
    d = ctx.get_samples()["disp"]
    for e in range(8):
        np.testing.assert_allclose(d[:, e, 0], want, rtol=0, atol=1e-12 * v * t_s)
    ctx.close()
    # a drain at the far end: uniform motion v t and floor(v t / L) drains
    flags[-1] = 2
    ctx, _ = _tube_context(kmc, n, a, v, flags)
    ctx.inject(8)
    ctx.set_samples(nb, t_s)
    ctx.run(10)
    smp = ctx.get_samples()
    u = v * tk
    fr = np.mod(u / L, 1.0)
    keep = (fr > 1e-6) & (fr < 1 - 1e-6)
    for e in range(8):
        np.testing.assert_allclose(smp["disp"][keep, e, 0], u[keep], rtol=1e-12, atol=0)
        assert np.array_equal(smp["drains"][keep, e], np.floor(u[keep] / L))
    ctx.close()


def test_velocity_errors(kmc, nw, small):
    say("velocity argument and geometry errors")
    ctx = _ctx(kmc, nw, small, 0.0)
    with pytest.raises(RuntimeError):
        ctx.set_velocity(-1.0)
    with pytest.raises(RuntimeError):
        ctx.set_velocity(float("inf"))
    # an ensemble member must hop without drift
    ctx.set_velocity(1.0)
    ctx.inject(64)
    with pytest.raises(RuntimeError, match="velocity"):
        kmc.Ensemble([ctx]).run(10)
    ctx.close()
    # two linked sites at one position
    n = 3
    pos = np.array([[0.0, 0, 0], [0.0, 0, 0], [10 * NM, 0, 0]])
    ori = np.tile([1.0, 0.0, 0.0], (n, 1))
    s = site_arrays(pos, ori, flags=np.array([1, 0, 0], np.uint8),
                    left=np.array([-1, 0, 1], np.int32), right=np.array([1, 2, -1], np.int32))
    c2 = kmc.Context(kmc.Config(cutoff=5 * NM, min_distance=0.4e-9, seed=1))
    c2.set_sites(s["x"], s["y"], s["z"], s["ox"], s["oy"], s["oz"], s["left"], s["right"],
                 s["chirality"], s["flags"])
    with pytest.raises(RuntimeError, match="coincide"):
        c2.set_velocity(1.0)
    c2.close()


def test_pass_bound_is_an_error(kmc):
    # 64 excitons on a 3-site tube asked to fly 1e8 site lengths: every flight stops at the
    # 2^22-pass bound (bounded work) and kmc_run reports it
    say("flight pass bound")
    flags = np.array([1, 0, 0], np.uint8)
    ctx, _ = _tube_context(kmc, 3, 10 * NM, 1.0, flags)
    ctx.inject(64)
    ctx.set_samples(1, 1.0)
    with pytest.raises(RuntimeError, match="2\\^22"):
        ctx.run(1)
    ctx.close()
