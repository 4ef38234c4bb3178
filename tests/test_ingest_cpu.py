"""Site ingest on the host: the mesh reader (create_scatterers, src/monte_carlo/monte_carlo.h:
199-271), trim_scats (:722-782) and the Kubo regions (injection_region, get_removal_domain,
src/monte_carlo/monte_carlo.cpp:203-251), through the C ABI (host-only entry points, no GPU),
against the oracle's literal restatement (oracle/ingest_oracle.py), which is pinned here by a
hand-worked trim, brute-force invariants, a grid whose regions are known in closed form and a
file round trip."""
import numpy as np
import pytest

from conftest import write_mesh

NM = 1e-9


@pytest.fixture(scope="module")
def ing():
    from oracle import ingest_oracle
    return ingest_oracle


def tiny_mesh(rows=3, cols=4, seed=5):
    rng = np.random.default_rng(seed)
    pos = [np.round(rng.uniform(-500, 500, (rows, cols)), 3) for _ in range(3)]
    ori = rng.normal(size=(3, rows, cols))
    ori /= np.linalg.norm(ori, axis=0)
    return pos, list(ori)


@pytest.mark.parametrize("sep", [" ", ",", "\t"])
def test_mesh_round_trip(kmc, ing, tmp_path, sep):
    pos, ori = tiny_mesh()
    write_mesh(tmp_path, pos, ori, sep)
    got = kmc.read_mesh(tmp_path)
    want = ing.mesh_sites(pos, ori)
    for k in ("x", "y", "z", "ox", "oy", "oz", "left", "right"):
        assert np.array_equal(got[k], want[k]), k
    # nm -> m by one FP64 multiply; links along each row, -1 at the tube ends
    assert np.array_equal(got["x"], pos[0].ravel() * 1.e-9)
    assert list(got["left"][:4]) == [-1, 0, 1, 2] and list(got["right"][:4]) == [1, 2, 3, -1]
    assert got["n"] == 12 and list(got["tube"]) == [0] * 4 + [1] * 4 + [2] * 4


def test_mesh_errors(kmc, tmp_path):
    pos, ori = tiny_mesh()
    with pytest.raises(kmc.KmcError, match="cannot open"):
        kmc.read_mesh(tmp_path)
    write_mesh(tmp_path, pos, ori)
    (tmp_path / "single_cnt.orient.z.dat").write_text("1 2 3 4\n5 6 7\n8 9 10 11\n")
    with pytest.raises(kmc.KmcError, match="ragged"):
        kmc.read_mesh(tmp_path)
    (tmp_path / "single_cnt.orient.z.dat").write_text("1 2 3 4\n5 6 7 8\n")
    with pytest.raises(kmc.KmcError, match="different shape"):
        kmc.read_mesh(tmp_path)
    (tmp_path / "single_cnt.orient.z.dat").write_text("1 2 3 4\n5 x 7 8\n9 9 9 9\n")
    with pytest.raises(kmc.KmcError, match="bad number"):
        kmc.read_mesh(tmp_path)


def _two_tubes():
    x = np.array([0.0, 9.0, 2.0, 3.0, 4.0, 9.0]) * NM   # sites 1 and 5 lie outside x <= 5 nm
    z0 = np.zeros(6)
    return dict(x=x, y=z0.copy(), z=z0.copy(), ox=np.ones(6), oy=z0.copy(), oz=z0.copy(),
                left=np.array([-1, 0, 1, -1, 3, 4], np.int32),
                right=np.array([1, 2, -1, 4, 5, -1], np.int32),
                flags=np.array([1, 0, 0, 0, 2, 0], np.uint8))


def test_trim_hand_worked(kmc, ing):
    # slot order and links worked out by hand from trim_scats' swap-with-last loop
    net = _two_tubes()
    lim = ((-1.0, 5 * NM), (-1.0, 1.0), (-1.0, 1.0))
    for got in (kmc.trim_sites(net, *lim), ing.trim(net, *lim)):
        origin = got["source_index"] if "source_index" in got else got["origin"]
        assert list(origin) == [0, 4, 2, 3]
        assert list(got["left"]) == [-1, 3, -1, -1]
        assert list(got["right"]) == [-1, -1, -1, 1]
        assert np.array_equal(got["x"], net["x"][[0, 4, 2, 3]])
    assert list(kmc.trim_sites(net, *lim)["flags"]) == [1, 2, 0, 0]


def _check_trim_invariants(net, got, lim):
    inside = ((net["x"] >= lim[0][0]) & (net["x"] <= lim[0][1]) & (net["y"] >= lim[1][0]) &
              (net["y"] <= lim[1][1]) & (net["z"] >= lim[2][0]) & (net["z"] <= lim[2][1]))
    org = got["source_index"]
    assert sorted(org) == sorted(np.nonzero(inside)[0])
    slot_of = {int(o): s for s, o in enumerate(org)}
    for s, o in enumerate(org):
        for side, other in (("left", "right"), ("right", "left")):
            t = got[side][s]
            if t >= 0:
                assert got[other][t] == s                          # links stay mutual
                assert net[side][o] == org[t]                      # and come from the mesh
            else:
                orig = net[side][o]
                assert orig < 0 or not inside[orig]                # dropped only to removed sites
    assert len(slot_of) == len(org)


def test_trim_film_matches_oracle_and_invariants(kmc, ing, nw):
    net = nw.build_network(nw.film_small())
    lim = ((100 * NM, 420 * NM), (-1.0, 1.0), (5 * NM, 45 * NM))
    got = kmc.trim_sites(net, *lim)
    want = ing.trim(net, *lim)
    for k in ("x", "y", "z", "ox", "oy", "oz", "left", "right"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["source_index"], want["origin"])
    assert 0 < got["n"] < net["x"].size
    _check_trim_invariants(net, got, lim)


def test_kubo_regions_on_a_grid(kmc, ing):
    g = np.arange(5, dtype=float) * 10 * NM
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    net = dict(x=X.ravel(), y=Y.ravel(), z=Z.ravel())
    for flags, lo, hi in (kmc.kubo_regions(net, 5), ing.kubo_regions(net["x"], net["y"],
                                                                     net["z"], 5)):
        # edges 0, 8, 16, 24, 32, 40 nm: the centre cell [16, 24] holds the grid point 20 nm;
        # the removal box [8, 32] nm excludes every site with a coordinate at 0 or 40 nm
        src = np.nonzero(flags & 1)[0]
        assert list(src) == [62]
        assert int(np.sum((flags & 2) != 0)) == 125 - 27
        assert np.array_equal(lo, [0, 0, 0]) and np.array_equal(hi, [40 * NM] * 3)
    with pytest.raises(kmc.KmcError, match="odd"):
        kmc.kubo_regions(net, 4)


def test_kubo_regions_film_matches_oracle(kmc, ing, nw):
    net = nw.build_network(nw.film_small())
    f, lo, hi = kmc.kubo_regions(net, 5)
    fo, loo, hio = ing.kubo_regions(net["x"], net["y"], net["z"], 5)
    assert np.array_equal(f, fo) and np.array_equal(lo, loo) and np.array_equal(hi, hio)
    assert np.any(f & 1) and np.any(f & 2)


def test_generated_film_through_mesh_files(kmc, nw, tmp_path):
    # the 5k-site film written as a mesh (one row per 20-site tube, nm) reads back to the same
    # sites and links
    net = nw.build_network(nw.film_small())
    npt = 20
    rows = net["x"].size // npt
    assert np.array_equal(net["tube"], np.repeat(np.arange(rows), npt))
    pos = [np.asarray(net[k]).reshape(rows, npt) * 1e9 for k in ("x", "y", "z")]
    ori = [np.asarray(net[k]).reshape(rows, npt) for k in ("ox", "oy", "oz")]
    write_mesh(tmp_path, pos, ori)
    got = kmc.read_mesh(tmp_path)
    assert np.array_equal(got["left"], net["left"]) and np.array_equal(got["right"], net["right"])
    for k in ("x", "y", "z"):
        np.testing.assert_allclose(got[k], net[k], rtol=4e-16, atol=1e-30)
    for k in ("ox", "oy", "oz"):
        assert np.array_equal(got[k], net[k])
