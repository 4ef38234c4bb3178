"""The reference's own set-up chain on the GPU: mesh files -> create_scatterers -> trim_scats ->
kubo_init regions (src/monte_carlo/monte_carlo.cpp:254-302) -> cell list, CSR and hops, no
periodic axis (the reference has none), checked bit-exactly against the oracle on the same
trimmed sites."""
import sys

import numpy as np
import pytest

from conftest import write_mesh

pytestmark = pytest.mark.gpu
NM = 1e-9


def say(*a):
    print("[gpu-ingest]", *a, file=sys.stderr, flush=True)


def test_mesh_trim_regions_to_hops(kmc, orc, nw, tmp_path):
    net0 = nw.build_network(nw.film_small())
    npt = 20
    rows = net0["x"].size // npt
    write_mesh(tmp_path, [np.asarray(net0[k]).reshape(rows, npt) * 1e9 for k in ("x", "y", "z")],
               [np.asarray(net0[k]).reshape(rows, npt) for k in ("ox", "oy", "oz")])
    mesh = kmc.read_mesh(tmp_path)
    net = kmc.trim_sites(mesh, (50 * NM, 450 * NM), (20 * NM, 480 * NM), (-1.0, 1.0))
    flags, lo, hi = kmc.kubo_regions(net, 5)
    net["flags"] = flags
    net["chirality"] = np.zeros(net["n"], np.uint8)
    say(f"mesh {mesh['n']} sites -> trimmed {net['n']}; sources {int(np.sum(flags & 1))}, "
        f"drain sites {int(np.sum((flags & 2) != 0))}")
    cutoff = 15 * NM
    th, z, a1, a2, rates = nw.forster_tables(1)
    ctx = kmc.Context(kmc.Config(cutoff=cutoff, min_distance=0.4 * NM, seed=23))
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"], net["left"],
                  net["right"], net["chirality"], net["flags"])
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    csr = ctx.get_csr()
    onet = orc.Net(net, cutoff, 0.4 * NM)
    onet.set_table(th, z, a1, a2, rates)
    ocsr = onet.csr(0)
    assert np.array_equal(csr["row_ptr"], ocsr["row_ptr"])
    assert np.array_equal(csr["col"], ocsr["col"]) and np.array_equal(csr["cum"], ocsr["cum"])
    E, H = 512, 2000
    ctx.inject(E)
    ctx.run(H)
    g = ctx.get_excitons()
    ox = orc.Excitons(csr, net["flags"], seed=23, n_excitons=E)
    ox.run(H)
    assert np.array_equal(g["site"], ox.site) and np.array_equal(g["hash"], ox.hash)
    assert np.array_equal(g["drains"], ox.drains) and np.array_equal(g["disp"], ox.dsp)
    assert np.array_equal(ctx.get_hop_counts(), ox.site_count)
    assert int(g["drains"].sum()) > 0
    np.testing.assert_allclose(g["t_next"], ox.t_next, rtol=1e-12, atol=0)
    ctx.close()
