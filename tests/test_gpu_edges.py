"""Edge cases of the hop path on the GPU against the CPU oracle: empty inputs, exciton counts
that fill no wave or workgroup evenly, a zero hop budget, a time cap below every first dwell,
and rows far wider than one 128-byte line (sites with 50 to 187 neighbours, whose cumulative
rates and targets continue over up to 18 lines).  Bar as in the other parity tests: sites, hop
counts, visit hashes, drains, displacements and per-site counts bitwise, times within 1e-12
relative (update_state and ff_time, scatterer.cpp:9-31, scatterer.h:74-80)."""
import sys

import numpy as np
import pytest

from conftest import site_arrays

pytestmark = pytest.mark.gpu
NM = 1e-9
T_REL = 1e-12
BIG = (1 << 63) // 2


def say(*a):
    print("[gpu-edges]", *a, file=sys.stderr, flush=True)


def _compare_state(g, o, label):
    assert np.array_equal(g["site"], o.site), label
    assert np.array_equal(g["hops"], o.hops), label
    assert np.array_equal(g["hash"], o.hash), label
    assert np.array_equal(g["drains"], o.drains), label
    assert np.array_equal(g["disp"], o.dsp), label
    fin = np.isfinite(o.t_next)
    assert np.array_equal(np.isfinite(g["t_next"]), fin), label
    np.testing.assert_allclose(g["t_next"][fin], o.t_next[fin], rtol=T_REL, atol=0)


@pytest.fixture(scope="module")
def film(kmc, orc, nw):
    w = nw.film_small()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(1)
    ctx = nw.make_context(w, net, seed=31)
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    csr = ctx.get_csr()
    yield dict(w=w, net=net, ctx=ctx, csr=csr)
    ctx.close()


def test_empty_inputs_are_errors(kmc, film):
    say("empty inputs")
    ctx = kmc.Context(kmc.Config(cutoff=15 * NM, seed=1))
    z = np.zeros(0)
    with pytest.raises(kmc.KmcError) as e:
        ctx.set_sites(z, z, z, z, z, z)
    assert e.value.code == kmc.ERR_ARG
    ctx.close()
    c = film["ctx"]
    with pytest.raises(kmc.KmcError) as e:
        c.inject(0)
    assert e.value.code == kmc.ERR_ARG
    c.inject(100)
    with pytest.raises(kmc.KmcError) as e:
        c.run(-1)
    assert e.value.code == kmc.ERR_ARG


@pytest.mark.parametrize("E", [1, 63, 65, 257, 1023])
def test_ragged_exciton_counts_match_oracle(kmc, orc, film, E):
    """E excitons that leave the last wave or workgroup partly empty: the idle lanes take part
    in every cooperative row load and count nothing."""
    H = 3000
    say(f"ragged: {E} excitons x {H} hops")
    c = film["ctx"]
    c.inject(E)
    ox = orc.Excitons(film["csr"], film["net"]["flags"], seed=31, n_excitons=E)
    _compare_state(c.get_excitons(), ox, "after injection")
    c.run(H)
    ox.run(H)
    _compare_state(c.get_excitons(), ox, f"E={E}")
    assert np.array_equal(c.get_hop_counts(), ox.site_count)
    assert int(ox.site_count.sum()) == E * H


def test_zero_budget_and_early_cap(kmc, orc, film):
    """A zero hop budget changes nothing; a time cap below every exciton's first dwell makes no
    hop and leaves every clock as it was; the run after them equals the oracle's."""
    say("zero budget and early cap")
    c = film["ctx"]
    E = 500
    c.inject(E)
    g0 = c.get_excitons()
    c.run(0)
    g1 = c.get_excitons()
    for k in g0:
        assert np.array_equal(g0[k], g1[k]), k
    t_cap = 0.5 * float(np.min(g0["t_next"]))
    c.run(BIG, t_max=t_cap)
    g2 = c.get_excitons()
    for k in g0:
        assert np.array_equal(g0[k], g2[k]), k
    assert int(c.get_hop_counts().sum()) == 0
    # no exciton can hop before the cap, so the run stopped after its first launch
    assert c.run_stats() == (1, 0)
    ox = orc.Excitons(film["csr"], film["net"]["flags"], seed=31, n_excitons=E)
    c.run(700)
    ox.run(700)
    _compare_state(c.get_excitons(), ox, "after the early cap")


def test_wide_rows_match_oracle(kmc, orc, nw):
    """A 6 x 6 x 6 grid of segments 4 nm apart with random orientations and a 15 nm cutoff:
    every site has 50 to 187 neighbours, so every row continues beyond its first line (5 to 18
    lines of 128 bytes).  CSR and hops (selections and targets read from continuation lines)
    bitwise against the oracle."""
    rng = np.random.default_rng(17)
    g = np.arange(6) * 4 * NM
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    ori = rng.normal(size=pos.shape)
    ori /= np.linalg.norm(ori, axis=1, keepdims=True)
    flags = np.zeros(len(pos), np.uint8)
    flags[pos[:, 0] == 0] = 1                       # sources on the x = 0 face
    flags[pos[:, 0] == g[-1]] |= 2                  # drains on the far face
    net = site_arrays(pos, ori, flags=flags)
    th, z, a1, a2, rates = nw.forster_tables(1)
    onet = orc.Net(net, 15 * NM, 0.4 * NM, 0)
    onet.set_table(th, z, a1, a2, rates)
    ocsr = onet.csr(0)
    deg = np.diff(ocsr["row_ptr"])
    say(f"wide rows: {len(pos)} sites, degree {deg.min()}..{deg.max()}")
    assert deg.min() > 8 and deg.max() > 150         # every row has continuation lines
    ctx = kmc.Context(kmc.Config(cutoff=15 * NM, seed=5))
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"], flags=flags)
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    gcsr = ctx.get_csr()
    for k in ("row_ptr", "col", "cum", "disp"):
        assert np.array_equal(gcsr[k], ocsr[k]), k
    E, H = 2000, 4000
    ctx.inject(E)
    ox = orc.Excitons(gcsr, flags, seed=5, n_excitons=E)
    ctx.run(H)
    ox.run(H)
    _compare_state(ctx.get_excitons(), ox, "wide rows")
    assert np.array_equal(ctx.get_hop_counts(), ox.site_count)
    assert int(ox.drains.sum()) > 0
    ctx.close()
