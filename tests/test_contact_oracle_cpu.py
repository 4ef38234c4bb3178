"""The oracle's contact mode (orc_contact: init, step, save_metrics, repopulate_contacts of
src/monte_carlo/monte_carlo.h, restated literally) pinned without the GPU: section areas of
get_area on a lattice against the closed form, the initial linear profile of create_particles
against round(c1 + i (c2 - c1) / (n_seg - 1)), particle conservation across steps, and Fick's law
(visualization/profiler.py) against the known D = k a^2 of a nearest-neighbour lattice."""
import numpy as np

NM = 1e-9


def _lattice(orc, nx, ny, nz, a, k):
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    pos = g.reshape(-1, 3).astype(float) * a
    n = pos.shape[0]
    net = dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(),
               ox=np.zeros(n), oy=np.ones(n), oz=np.zeros(n), left=np.full(n, -1, np.int32),
               right=np.full(n, -1, np.int32), chirality=np.zeros(n, np.uint8),
               flags=np.zeros(n, np.uint8), n=n)
    w = orc.Net(net, 1.2 * a, 0.4 * NM)
    w.set_table([0.0], [0.0], [0.0], [0.0], np.full((1, 1, 1, 1), k))
    return net, w.csr(0)


def test_areas_profile_and_conservation(orc):
    a, k = 10 * NM, 1e12
    nx, ny, nz = 3, 40, 5
    net, csr = _lattice(orc, nx, ny, nz, a, k)
    assert np.all(np.diff(csr["row_ptr"]) <= 6)
    n_seg, c1, c2 = 5, 200, 30
    # one step too short for any hop: the population is the initial profile
    r = orc.contact(net, csr, 3, n_seg, c1, c2, 5000, 1e-20, 1)
    # get_area: x extent times z extent of each section's sites
    np.testing.assert_allclose(r["area"], (nx - 1) * a * (nz - 1) * a, rtol=1e-15)
    want = [int(np.floor(c1 + i * (c2 - c1) / (n_seg - 1) + 0.5)) for i in range(n_seg)]
    assert list(r["population"][0]) == want
    assert np.all(r["crossings"] == 0) and int(r["id"].size) == r["n_live"][0]
    # conservation: a step's population is the previous step's live count; the repopulation
    # keeps everything outside the contacts and adds c1 + c2
    r = orc.contact(net, csr, 3, n_seg, c1, c2, 5000, 2e-12, 300)
    pop, live = r["population"], r["n_live"]
    assert pop[0].sum() == sum(want)
    assert np.array_equal(pop[1:].sum(1), live[:-1])
    assert np.array_equal(live, pop[:, 1:-1].sum(1) + c1 + c2)
    assert np.all(np.diff(np.sort(r["id"])) > 0)              # ids stay unique
    assert r["hops_made"] >= int(r["hops"].sum()) > 0         # removed particles' hops too
    assert int(r["id"].max()) < sum(want) + 300 * (c1 + c2)


def test_fick_known_diffusion_oracle(orc, kmc):
    a, k = 10 * NM, 1e12
    nx, ny, nz = 4, 30, 4
    net, csr = _lattice(orc, nx, ny, nz, a, k)
    dt = 1e-12
    r = orc.contact(net, csr, 17, 5, 300, 0, 20000, dt, 8000)
    ymin, dy = 0.0, (ny - 1) * a / 5
    D = kmc.fick_diffusion(r["population"], r["crossings"], r["area"], ymin, dy, dt,
                           steady_from=2000, interior_fit=True)[0]
    # measured -6.6 %: the reference's sections are (ny - 1) a / n_seg = 5.8 a long but hold
    # ny / n_seg = 6 layers each, so the section centres the gradient uses are 3.4 % closer than
    # the layers they average; the rest is the step-wise contact reset and noise
    assert abs(D / (k * a * a) - 1) < 0.1


def test_drift_vanishing_velocity_is_hopping(orc, kmc):
    # v -> 0: no flight reaches the next site, so the run is the hopping run with y unchanged
    a, k = 10 * NM, 1e12
    nx, ny, nz = 3, 30, 3
    net, csr = _lattice(orc, nx, ny, nz, a, k)
    _tubes_along_y(net, ny, nz)
    r0 = orc.contact(net, csr, 5, 5, 120, 10, 5000, 2e-12, 60)
    r1 = orc.contact(net, csr, 5, 5, 120, 10, 5000, 2e-12, 60, velocity=1e-30)
    for key in ("population", "crossings", "n_live", "id", "site", "hops", "t_next", "y"):
        assert np.array_equal(r0[key], r1[key]), key


def _tubes_along_y(net, ny, nz):
    # every (x, z) column a tube along y: right = the next site up, left = the one down
    idx = np.arange(net["n"])
    iy = (idx // nz) % ny
    net["right"] = np.where(iy < ny - 1, idx + nz, -1).astype(np.int32)
    net["left"] = np.where(iy > 0, idx - nz, -1).astype(np.int32)


def test_drift_pure_flight_closed_form(orc):
    """No rates: every particle only flies (particle::fly, particle.cpp:9-54) along its straight
    tube at v, turning round at the tube ends, so y(t) = fold(y0 + s v t) with fold the
    reflection into [0, L].  A tiny first step gives each particle's start y0 and heading s; a
    second run from the same seed, one step of 2.3 a / v, must put it at fold(y0 + 2.3 s a)."""
    a, v = 10 * NM, 1.0
    nx, ny, nz = 3, 40, 3
    net, csr = _lattice(orc, nx, ny, nz, a, 0.0)
    _tubes_along_y(net, ny, nz)
    L = (ny - 1) * a
    tiny = orc.contact(net, csr, 9, 4, 50, 50, 2000, 0.01 * a / v, 1, velocity=v)
    off = tiny["y"] - net["y"][tiny["site"]]
    assert np.all(np.isinf(tiny["t_next"]))
    n0 = int(tiny["population"][0].sum())                    # the initial particles: ids < n0
    first = tiny["id"] < n0
    assert np.allclose(np.abs(off[first]), 0.01 * a, rtol=1e-5)
    s = np.sign(off[first])
    assert 0.3 < np.mean(s > 0) < 0.7                         # both headings drawn
    y0 = dict(zip(tiny["id"][first].tolist(), zip(net["y"][tiny["site"][first]], s)))
    far = orc.contact(net, csr, 9, 4, 50, 50, 2000, 2.3 * a / v, 1, velocity=v)
    checked = 0
    for pid, y in zip(far["id"].tolist(), far["y"]):
        if pid >= n0 or pid not in y0:
            continue
        u = (y0[pid][0] + y0[pid][1] * 2.3 * a) % (2 * L)
        want = L - abs(u - L)
        assert abs(y - want) < 1e-4 * a, (pid, y, want)
        checked += 1
    assert checked > 40
