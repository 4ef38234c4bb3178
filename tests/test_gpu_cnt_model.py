"""The CNT exciton model with its two large sums on the GPU (kmc_cnt_create_on_device: the Ohno
transform v(q), cnt.cpp:520-631, and the polarization PI(q), cnt.cpp:634-732) against the
host build (kmc_cnt_create) and the oracle's numpy restatement.  The device sums run in
another order with the device sincos / hypot, so the bar is rounding: v(q) and PI(q) within
1e-13 of each array's scale, exciton energies within 1e-12 relative, transfer rates built on
the two models within 1e-9 relative."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-cnt]", *a, file=sys.stderr, flush=True)


CHIRALITIES = [(4, 2), (7, 5), (8, 6), (9, 7)]


@pytest.fixture(scope="module")
def pairs():
    from cnt_film_monte_carlo_amd import exciton as X
    out = {}
    for c in CHIRALITIES:
        t = time.perf_counter()
        host = X.CntModel(*c, 10)
        t_host = time.perf_counter() - t
        t = time.perf_counter()
        dev = X.CntModel(*c, 10, device=0)
        t_dev = time.perf_counter() - t
        say(f"{c}: host build {t_host:.3f} s, with the GPU sums {t_dev:.3f} s")
        out[c] = (host, dev)
    return out


@pytest.mark.parametrize("c", CHIRALITIES)
def test_device_sums_match_host_model(pairs, c):
    host, dev = pairs[c]
    for f in ("nk", "n_mu", "nk_relev", "nk_cm", "Nu"):
        assert getattr(host.info, f) == getattr(dev.info, f)
    vh, ph, eh = host.interaction()
    vd, pd, ed = dev.interaction()
    scale = np.abs(vh).max(axis=(0, 1))
    err = np.abs(vd - vh).max(axis=(0, 1)) / scale
    assert np.all(err < 1e-13), err
    assert np.abs(pd - ph).max() < 1e-13 * ph.max()
    assert np.abs(ed - eh).max() < 1e-12 * np.abs(eh).max()
    np.testing.assert_allclose(dev.dispersion(), host.dispersion(), rtol=1e-12, atol=0)
    say(f"{c}: v(q) {err.max():.1e}, PI {np.abs(pd - ph).max() / ph.max():.1e} of scale")


@pytest.mark.parametrize("c", [(4, 2), (7, 5)])
def test_device_sums_match_numpy(pairs, c):
    from oracle import cnt_oracle as co
    _, dev = pairs[c]
    i = dev.info
    pa, pb, _ = dev.atoms()
    e, wf = dev.electrons()
    vd, pd, _ = dev.interaction()
    ov = co.vq(i, pa, pb)
    op = co.polarization(i, e, wf)
    assert np.all(np.abs(vd - ov).max(axis=(0, 1)) < 1e-13 * np.abs(ov).max(axis=(0, 1)))
    assert np.abs(pd - op).max() < 1e-13 * op.max()


def test_rates_on_device_built_models(pairs):
    from cnt_film_monte_carlo_amd import exciton as X
    g = np.array([(1.5e-9, 0, 0, 0.0), (4e-9, -3e-9, 2e-9, 0.7), (2.5e-9, 1e-9, 0, 2.0),
                  (7e-9, 2e-9, -5e-9, 1.5707963)])
    for d, a in [((7, 5), (7, 5)), ((8, 6), (7, 5)), ((7, 5), (9, 7))]:
        th = X.Transfer(pairs[d][0], pairs[a][0], 300.0)
        td = X.Transfer(pairs[d][1], pairs[a][1], 300.0)
        assert th.info.n_terms == td.info.n_terms
        rh = th.rates(g[:, 3], g[:, 0], g[:, 1], g[:, 2])
        rd = td.rates(g[:, 3], g[:, 0], g[:, 1], g[:, 2])
        np.testing.assert_allclose(rd, rh, rtol=1e-9, atol=0)
        th.close()
        td.close()


def test_bad_device_is_an_error():
    from cnt_film_monte_carlo_amd import exciton as X
    from cnt_film_monte_carlo_amd import kmc
    with pytest.raises(kmc.KmcError, match="device"):
        X.CntModel(4, 2, 10, device=-1)
    with pytest.raises(kmc.KmcError, match="no device"):
        X.CntModel(4, 2, 10, device=4096)
