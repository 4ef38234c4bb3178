"""Contact (steady-state) mode on the GPU: init / step / save_metrics / repopulate_contacts
(src/monte_carlo/monte_carlo.h:157-195, 274-316, 343-355, 443-684) against the oracle's literal
restatement (orc_contact) bit for bit -- populations, interface crossings, live counts, section
areas, particles -- and Fick's law D = J / |dn/dy| (visualization/profiler.py:25-65) on a
simple cubic lattice, where D = k a^2 is known."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
NM = 1e-9


def say(*a):
    print("[gpu-contact]", *a, file=sys.stderr, flush=True)


def _film(kmc, nw, seed, velocity=0.0, open_tubes=False):
    # the 5k-site film with open axes (transport along y) and no Kubo flags
    w = nw.film_small()
    net = nw.build_network(w)
    net["flags"] = np.zeros(net["n"], np.uint8)
    if open_tubes:
        # the tubes wrap around the film's y period; with open axes that link is a ~500 nm
        # segment across the film, so the drift tests cut it and the tubes end at the edges
        for side, other in (("left", "right"), ("right", "left")):
            lk = net[side]
            far = (lk >= 0) & (np.abs(net["y"][np.maximum(lk, 0)] - net["y"]) > 0.25 * np.ptp(net["y"]))
            net[side] = np.where(far, -1, lk).astype(np.int32)
    ctx = kmc.Context(kmc.Config(cutoff=w.cutoff, min_distance=0.4 * NM, seed=seed))
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"], net["left"],
                  net["right"], net["chirality"], net["flags"])
    tab = nw.forster_tables(1)
    ctx.set_rate_table(*tab)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    if velocity > 0.0:
        ctx.set_velocity(velocity)
    return w, net, ctx, tab


# the third case makes ~170 hops per particle and step, so from its second step on every step
# runs as capped passes (pauses, resumed particles, the caps carried from one call to the next)
@pytest.mark.parametrize("n_seg,c1,c2,dt,steps", [(10, 400, 40, 2e-8, 80), (3, 150, 0, 5e-8, 40),
                                                  (10, 400, 40, 4e-7, 24)])
def test_contact_matches_oracle(kmc, orc, nw, n_seg, c1, c2, dt, steps):
    say(f"contact mode vs oracle: n_seg {n_seg}, populations {c1}/{c2}, {steps} steps of {dt}")
    w, net, ctx, tab = _film(kmc, nw, seed=29)
    cap = 8000
    ctx.contact_init(n_seg, c1, c2, cap)
    p0 = ctx.contact_particles()
    pop, cur, live = ctx.contact_run(dt, steps // 2)
    pop2, cur2, live2 = ctx.contact_run(dt, steps - steps // 2)   # the run continues
    pop, cur, live = np.vstack([pop, pop2]), np.vstack([cur, cur2]), np.concatenate([live, live2])
    onet = orc.Net(net, w.cutoff, 0.4 * NM)
    onet.set_table(*tab)
    ref = orc.contact(net, onet.csr(0), 29, n_seg, c1, c2, cap, dt, steps)
    area, ymin, dy, T = ctx.contact_sections()
    assert np.array_equal(area, ref["area"])
    assert T == pytest.approx(steps * dt, rel=1e-12)
    assert np.array_equal(pop, ref["population"])
    assert np.array_equal(cur, ref["crossings"])
    assert np.array_equal(live, ref["n_live"])
    p = ctx.contact_particles()
    assert np.array_equal(p["id"], ref["id"]) and np.array_equal(p["site"], ref["site"])
    assert np.array_equal(p["hops"], ref["hops"])
    np.testing.assert_allclose(p["t_next"], ref["t_next"], rtol=1e-12)
    # every hop made, those of particles since removed included, and the step kernels' time
    made, step_ms = ctx.contact_stats()
    assert made == ref["hops_made"] and step_ms > 0.0
    # bookkeeping: a step's population counts the particles that entered it
    assert pop[0].sum() == p0["id"].size
    assert np.array_equal(pop[1:].sum(1), live[:-1])
    assert int(p["hops"].sum()) > 0 and np.any(cur != 0)
    say(f"  live {live[-1]}, hops {int(p['hops'].sum())}, crossings {int(np.abs(cur).sum())}")
    ctx.close()


@pytest.mark.parametrize("v", [1.0, 10.0])
def test_contact_drift_matches_oracle(kmc, orc, nw, v):
    """v > 0: particle::fly (particle.cpp:9-54) to each hop and to every step's end, the
    particle's y = its site's y plus the offset along the segment toward the next site."""
    n_seg, c1, c2, dt, steps = 10, 400, 40, 2e-8, 80
    say(f"contact mode with drift v={v} m/s vs oracle: {steps} steps of {dt}")
    w, net, ctx, tab = _film(kmc, nw, seed=31, velocity=v, open_tubes=True)
    cap = 8000
    ctx.contact_init(n_seg, c1, c2, cap)
    pop, cur, live = ctx.contact_run(dt, steps // 2)
    pop2, cur2, live2 = ctx.contact_run(dt, steps - steps // 2)
    pop, cur, live = np.vstack([pop, pop2]), np.vstack([cur, cur2]), np.concatenate([live, live2])
    onet = orc.Net(net, w.cutoff, 0.4 * NM)
    onet.set_table(*tab)
    ref = orc.contact(net, onet.csr(0), 31, n_seg, c1, c2, cap, dt, steps, velocity=v)
    assert np.array_equal(pop, ref["population"])
    assert np.array_equal(cur, ref["crossings"])
    assert np.array_equal(live, ref["n_live"])
    p = ctx.contact_particles()
    assert np.array_equal(p["id"], ref["id"]) and np.array_equal(p["site"], ref["site"])
    assert np.array_equal(p["hops"], ref["hops"])
    # times (and so the flights' lengths) agree to rounding: the dwell times' logarithms are
    # the device's and libm's, equal to within an ulp or two
    np.testing.assert_allclose(p["y"], ref["y"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(p["t_next"], ref["t_next"], rtol=1e-12)
    assert ctx.contact_stats()[0] == ref["hops_made"]
    off = p["y"] - net["y"][p["site"]]
    assert np.mean(off != 0) > 0.5 and np.abs(off).max() <= 10.01 * NM   # between sites
    say(f"  live {live[-1]}, hops {int(p['hops'].sum())}, crossings {int(np.abs(cur).sum())}")
    ctx.close()


def test_contact_drift_refuses_drain_sites(kmc, nw):
    say("contact mode with drift refuses Kubo drain flags")
    w, net, ctx, tab = _film(kmc, nw, seed=3)
    ctx.close()
    net["flags"][:10] = kmc.SITE_DRAIN
    ctx = kmc.Context(kmc.Config(cutoff=w.cutoff, min_distance=0.4 * NM, seed=3))
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"], net["left"],
                  net["right"], net["chirality"], net["flags"])
    ctx.set_rate_table(*tab)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    ctx.contact_init(4, 100, 0, 300)                 # v = 0: drain flags play no part
    ctx.set_velocity(1.0)
    with pytest.raises(RuntimeError, match="KMC_SITE_DRAIN"):
        ctx.contact_init(4, 100, 0, 300)
    ctx.close()


def test_contact_errors(kmc, nw):
    w, net, ctx, tab = _film(kmc, nw, seed=5)
    with pytest.raises(kmc.KmcError) as e:
        ctx.contact_nseg = 4
        ctx.contact_run(1e-8, 1)
    assert e.value.code == kmc.ERR_STATE
    with pytest.raises(kmc.KmcError) as e:
        ctx.contact_init(1, 10, 0, 100)
    assert e.value.code == kmc.ERR_ARG
    with pytest.raises(kmc.KmcError) as e:
        ctx.contact_init(4, 1000, 0, 500)              # capacity below the initial profile
    assert e.value.code == kmc.ERR_CAPACITY
    ctx.contact_init(4, 100, 0, 300)
    with pytest.raises(kmc.KmcError, match="max_hops"):
        ctx.contact_run(1e-6, 1, max_hops_per_step=2)  # long step, tiny hop cap
    # the next call starts clean; the particles stopped by the cap catch up first (up to
    # ~1e6 hops each over the microsecond they are behind), hence the large cap
    pop, cur, live = ctx.contact_run(1e-12, 2, max_hops_per_step=1 << 30)
    assert pop.shape == (2, 4) and live[-1] > 0
    ctx.close()


def test_contact_near_capacity(kmc, orc, nw):
    """Both contacts held at 100 particles in 4 sections: with a capacity of 410 the live count
    reaches 408 and the run equals the oracle bit for bit (the steps' slot range is bounded by
    the live count and clamped at the capacity); with 400, the injections overflow it, which the
    oracle reports too, and the call fails with ERR_CAPACITY."""
    say("contact mode near the capacity: 4 sections, 100/100, capacities 410 and 400")
    w, net, ctx, tab = _film(kmc, nw, seed=29)
    ctx.contact_init(4, 100, 100, 410)
    pop, cur, live = ctx.contact_run(2e-8, 6)
    onet = orc.Net(net, w.cutoff, 0.4 * NM)
    onet.set_table(*tab)
    ref = orc.contact(net, onet.csr(0), 29, 4, 100, 100, 410, 2e-8, 6)
    assert np.array_equal(live, ref["n_live"]) and live.max() == 408
    assert np.array_equal(pop, ref["population"]) and np.array_equal(cur, ref["crossings"])
    p = ctx.contact_particles()
    assert np.array_equal(p["id"], ref["id"]) and np.array_equal(p["site"], ref["site"])
    assert np.array_equal(p["hops"], ref["hops"])
    ctx.close()
    with pytest.raises(RuntimeError):
        orc.contact(net, onet.csr(0), 29, 4, 100, 100, 400, 2e-8, 6)
    w, net, ctx, tab = _film(kmc, nw, seed=29)
    ctx.contact_init(4, 100, 100, 400)
    with pytest.raises(kmc.KmcError) as e:
        ctx.contact_run(2e-8, 6)
    assert e.value.code == kmc.ERR_CAPACITY and "capacity" in str(e.value)
    ctx.close()


def test_fick_known_diffusion(kmc):
    # simple cubic lattice, spacing a, cutoff 1.2 a (nearest neighbours only, the face
    # diagonals at 1.41 a excluded), rate k to each: along y the walk has D = k a^2
    a, k = 10 * NM, 1e12
    nx, ny, nz = 4, 60, 4
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    pos = g.reshape(-1, 3).astype(float) * a
    ori = np.tile([0.0, 1.0, 0.0], (pos.shape[0], 1))
    ctx = kmc.Context(kmc.Config(cutoff=1.2 * a, seed=11))
    ctx.set_sites(pos[:, 0], pos[:, 1], pos[:, 2], ori[:, 0], ori[:, 1], ori[:, 2])
    ctx.set_rate_table([0.0], [0.0], [0.0], [0.0], np.full(1, k))
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    assert ctx.nnz == 6 * nx * ny * nz - 2 * (ny * nz + nx * nz + nx * ny)
    ctx.contact_init(6, 1000, 0, 20000)
    dt = 1e-12                                         # ~2 hops per particle and step
    ctx.contact_run(dt, 6000)                          # to the steady state (L^2 / D = 3.5e-9 s)
    pop, cur, live = ctx.contact_run(dt, 8000)
    area, ymin, dy, _ = ctx.contact_sections()
    D, J, grad, n = kmc.fick_diffusion(pop, cur, area, ymin, dy, dt, interior_fit=True)
    D_ends = kmc.fick_diffusion(pop, cur, area, ymin, dy, dt)[0]
    say(f"  Fick D = {D:.4e} (interior fit), {D_ends:.4e} (end sections); k a^2 = {k * a * a:.4e}")
    say(f"  density profile {np.array2string(n, precision=3)}")
    # measured -3.6 % (fit) and +16 % (end sections): the sections are 59 a / 6 long but hold
    # 10 layers, so the gradient over section centres is 1.7 % steep; ~2 net crossings per step
    # over 8,000 steps leave ~1.5 % noise; the step-wise reset of the contacts bends the profile
    # next to them, which the end-section gradient of the reference's profiler takes in
    assert abs(D / (k * a * a) - 1) < 0.08
    assert abs(D_ends / (k * a * a) - 1) < 0.25
    ctx.close()


def test_contact_full_size_matches_oracle(kmc, orc, nw):
    """The 625k-site film with three chiralities at the bench's populations (20 sections,
    100,000 particles held in contact 1, ~1 M live) for three short steps (5 mean dwell times,
    ~30 hops per particle and step): populations, crossings, live counts, every particle and the
    hop total equal the oracle's.  This runs the full-size slot sort by region, chirality and row,
    the live-count bound and the wave kernel's XCD dealing against the literal restatement."""
    say("contact mode, 625k-site film vs oracle: 3 steps of 5 mean dwell times")
    w = nw.film_large()
    net = nw.build_network(w)
    net["flags"] = np.zeros(net["n"], np.uint8)
    tab = nw.forster_tables(3)
    ctx = kmc.Context(kmc.Config(cutoff=w.cutoff, min_distance=0.4 * NM, seed=77))
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"], net["left"],
                  net["right"], net["chirality"], net["flags"])
    ctx.set_rate_table(*tab)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    g = ctx.get_csr()
    gam = g["cum"][g["row_ptr"][1:][np.diff(g["row_ptr"]) > 0] - 1]
    dt = 5.0 * float(np.mean(1.0 / gam))
    ctx.contact_init(20, 100_000, 0, 4_000_000)
    pop, cur, live = ctx.contact_run(dt, 3)
    onet = orc.Net(net, w.cutoff, 0.4 * NM)
    onet.set_table(*tab)
    ref = orc.contact(net, onet.csr(0), 77, 20, 100_000, 0, 4_000_000, dt, 3)
    assert np.array_equal(pop, ref["population"])
    assert np.array_equal(cur, ref["crossings"])
    assert np.array_equal(live, ref["n_live"])
    p = ctx.contact_particles()
    assert np.array_equal(p["id"], ref["id"]) and np.array_equal(p["site"], ref["site"])
    assert np.array_equal(p["hops"], ref["hops"])
    np.testing.assert_allclose(p["t_next"], ref["t_next"], rtol=1e-12)
    made, _ = ctx.contact_stats()
    assert made == ref["hops_made"]
    say(f"  live {live[-1]}, hops {made}")
    ctx.close()


def test_contact_full_size_split_equals_whole(kmc, nw):
    """The 625k-site film at the bench's contact workload (20 sections, 100,000 particles held in
    contact 1, ~1 M live): a run split in two calls equals one call bit for bit (populations,
    crossings, live counts, every particle), the population of each step counts the particles
    that entered it, and every hop is counted.  Steps of 100 mean dwell times (~550 hops per
    particle) run as capped passes from the second step on."""
    say("contact mode, 625k-site film: 6 steps whole vs 3 + 3")
    w = nw.film_large()
    net = nw.build_network(w)
    net["flags"] = np.zeros(net["n"], np.uint8)
    tab = nw.forster_tables(3)

    def make():
        ctx = kmc.Context(kmc.Config(cutoff=w.cutoff, min_distance=0.4 * NM, seed=77))
        ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"], net["left"],
                      net["right"], net["chirality"], net["flags"])
        ctx.set_rate_table(*tab)
        ctx.build_cells()
        ctx.build_rates(kmc.RATE_TABLE)
        ctx.contact_init(20, 100_000, 0, 4_000_000)
        return ctx

    g = make().get_csr()
    gam = g["cum"][g["row_ptr"][1:][np.diff(g["row_ptr"]) > 0] - 1]
    dt = 100.0 * float(np.mean(1.0 / gam))
    a = make()
    pa, ca, la = a.contact_run(dt, 6)
    b = make()
    p1, c1, l1 = b.contact_run(dt, 3)
    p2, c2, l2 = b.contact_run(dt, 3)
    assert np.array_equal(pa, np.vstack([p1, p2]))
    assert np.array_equal(ca, np.vstack([c1, c2]))
    assert np.array_equal(la, np.concatenate([l1, l2]))
    qa, qb = a.contact_particles(), b.contact_particles()
    for k in ("id", "site", "hops", "t_next", "y"):
        assert np.array_equal(qa[k], qb[k]), k
    assert np.array_equal(pa[1:].sum(1), la[:-1])
    assert np.array_equal(la, pa[:, 1:-1].sum(1) + 100_000)
    ha, hb = a.contact_stats()[0], b.contact_stats()[0]
    assert ha == hb and ha >= int(qa["hops"].sum()) > 0
    say(f"  live {la[-1]}, hops {ha}")
    a.close()
    b.close()
