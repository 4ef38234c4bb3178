"""GPU tests of the trained rate surrogate (cnt_film_monte_carlo_amd/data/rate_surrogate.npz)
against the exact first-order rate, and of what it gains over building rates without it.

Bars (from the surrogate's specification): per-pair relative error small (median < 1.5%), at
most 1% of pairs beyond the 5% tolerance once flagged pairs went to the exact routine; the
surrogate-built config-1 table at least 20x faster than the exact build on the GPU; the
diffusion coefficient of a KMC run on surrogate rates within 2% of the run on exact rates with
the same seeds.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
NM = 1e-9


def say(*a):
    print("[gpu-surrogate-quality]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def S():
    from cnt_film_monte_carlo_amd import surrogate
    return surrogate


@pytest.fixture(scope="module")
def models(S):
    from cnt_film_monte_carlo_amd import exciton as X
    return {c: X.CntModel(*c, 10) for c in S.CHIRALITIES}


@pytest.fixture(scope="module")
def sur(S):
    return S.RateSurrogate.load()


def exact_transfer(models, S, cd, ca, T, dom):
    from cnt_film_monte_carlo_amd import exciton as X
    t = X.Transfer(models[S.CHIRALITIES[cd]], models[S.CHIRALITIES[ca]], T)
    t.set_clamp(dom.box_lo, dom.box_hi)
    return t


def test_model_file(S, sur):
    rep = sur.meta["train"]
    say(f"model: trained on {rep['n_train']} pairs, held-out median {rep['val_median_rel']:.4f}, "
        f"p99 {rep['val_p99_rel']:.4f}")
    assert sur.rmap is not None and sur.rmap_dims[0] == 9
    assert rep["val_median_rel"] < 0.01 and rep["val_frac_over_tol_unflagged"] < 0.01


def test_fresh_pairs_against_exact(S, sur, models):
    rng = np.random.default_rng(424242)   # a seed the training never used
    errs, flagged, over = [], 0, 0
    n_tot = 0
    for cd in range(3):
        for ca in range(3):
            for T in rng.uniform(150, 400, 2):
                th, z, a1, a2 = S.sample_geometries(rng, 2000, sur.domain)
                t = exact_transfer(models, S, cd, ca, float(T), sur.domain)
                ex = t.rates(th, z, a1, a2)
                r, fl, _ = sur.rates(th, z, a1, a2, cd, ca, float(T))
                e = S.relative_error(np.log10(np.maximum(r, 1e-300)), np.log10(np.maximum(ex, 1e-300)))
                ok = fl == 0
                errs.append(e[ok])
                flagged += int((~ok).sum())
                over += int((e[ok] > 0.05).sum())
                n_tot += e.size
    e = np.concatenate(errs)
    say(f"fresh pairs: median {np.median(e):.4f}, p99 {np.quantile(e, 0.99):.4f}, flagged "
        f"{flagged / n_tot:.4f}, over 5% unflagged {over / n_tot:.4f}")
    assert np.median(e) < 0.015
    assert over / n_tot < 0.01


def test_config1_table_build_gain(S, sur, models):
    from cnt_film_monte_carlo_amd import kmc
    th, z, a1, a2 = kmc.default_grids()
    t = exact_transfer(models, S, 0, 0, 300.0, sur.domain)
    t.rates([0.1], [2e-9], [0.0], [0.0])          # device upload outside the timing
    sur.rates([0.1], [2e-9], [0.0], [0.0], 0, 0, 300.0)
    ex = t.table(th, z, a1, a2)
    ms_exact = t.last_ms
    tab, n_fb, ms_mlp = sur.table(0, 0, 300.0, t, th, z, a1, a2)
    e = S.relative_error(np.log10(tab), np.log10(ex))
    say(f"config-1 table ({ex.size} nodes): exact {ms_exact:.1f} ms, surrogate {ms_mlp:.2f} ms "
        f"({ms_exact / ms_mlp:.0f}x), {n_fb} nodes exact; error median {np.median(e):.4f}, "
        f"max {e.max():.4f}, over 5% {np.mean(e > 0.05):.4f}")
    assert ms_exact / ms_mlp >= 20
    assert np.median(e) < 0.015 and np.mean(e > 0.05) < 0.01


def _diffusion(ctx, kmc, E, H, t_s, bins=32):
    ctx.inject(E)
    ctx.set_samples(bins, t_s)
    ctx.run(H)
    red, _ = ctx.reduce_samples(32)
    return kmc.estimate_diffusion(red, t_s, bins // 4, bins)


def test_film_small_rates_and_diffusion(S, kmc, nw, sur, models):
    w = nw.film_small()
    net = nw.build_network(w)
    t = exact_transfer(models, S, 0, 0, 300.0, sur.domain)
    ctx_e = nw.make_context(w, net, seed=77)
    ctx_e.build_cells()
    ctx_e.build_rates(kmc.RATE_PAIR)
    ms_e = ctx_e.build_rates_exact([t])
    ctx_m = nw.make_context(w, net, seed=77)
    ctx_m.build_cells()
    ctx_m.build_rates(kmc.RATE_PAIR)
    ms_m, n_fb = ctx_m.build_rates_mlp(sur, [0], 300.0, [t])
    re, rm = ctx_e.get_pair_rates(), ctx_m.get_pair_rates()
    e = S.relative_error(np.log10(rm), np.log10(re))
    say(f"film_small pair rates ({re.size}): exact {ms_e:.1f} ms, surrogate {ms_m:.2f} ms, "
        f"{n_fb} exact; error median {np.median(e):.4f}, over 5% {np.mean(e > 0.05):.4f}")
    assert np.median(e) < 0.015 and np.mean(e > 0.05) < 0.01
    # diffusion with the same seeds on both rate sets
    E, H = 200_000, 4000
    ctx_e.inject(E)
    ctx_e.run(H)
    tn = ctx_e.get_excitons()["t_next"]
    t_s = 0.8 * float(np.min(tn[np.isfinite(tn)]))
    De, ce = _diffusion(ctx_e, kmc, E, H, t_s)
    Dm, cm = _diffusion(ctx_m, kmc, E, H, t_s)
    say(f"D exact {De} +- {ce}; D surrogate {Dm} +- {cm}")
    for k in (0, 1):   # in-plane components
        assert abs(Dm[k] - De[k]) <= 0.02 * De[k], k
    ctx_e.close()
    ctx_m.close()


def test_mixed_chirality_film(S, kmc, nw, sur, models):
    # film_small geometry with the three chiralities of the large film (ids 0, 1, 2)
    w = nw.film_small()
    w.net.n_chir = 3
    net = nw.build_network(w)
    assert set(np.unique(net["chirality"])) == {0, 1, 2}
    tr = [exact_transfer(models, S, i, j, 300.0, sur.domain) for i in range(3) for j in range(3)]
    ctx_e = nw.make_context(w, net, seed=5)
    ctx_e.build_cells()
    ctx_e.build_rates(kmc.RATE_PAIR)
    ms_e = ctx_e.build_rates_exact(tr)
    ctx_m = nw.make_context(w, net, seed=5)
    ctx_m.build_cells()
    ctx_m.build_rates(kmc.RATE_PAIR)
    ms_m, n_fb = ctx_m.build_rates_mlp(sur, [0, 1, 2], 300.0, tr)
    re, rm = ctx_e.get_pair_rates(), ctx_m.get_pair_rates()
    e = S.relative_error(np.log10(np.maximum(rm, 1e-300)), np.log10(np.maximum(re, 1e-300)))
    say(f"mixed film ({re.size} pairs): exact {ms_e:.1f} ms, surrogate {ms_m:.2f} ms, {n_fb} exact; "
        f"error median {np.median(e):.4f}, over 5% {np.mean(e > 0.05):.4f}")
    assert np.median(e) < 0.015 and np.mean(e > 0.05) < 0.01
    # an unknown chirality id goes to the exact routine in full
    ctx_m.build_rates(kmc.RATE_PAIR)
    ms_u, n_u = ctx_m.build_rates_mlp(sur, [0, -1, 2], 300.0, tr)
    cls = ctx_m.get_pair_geometry()
    n_unknown = int(np.sum((cls["chir_src"] == 1) | (cls["chir_dst"] == 1)))
    assert n_u >= n_unknown
    np.testing.assert_array_equal(ctx_m.get_pair_rates()[(cls["chir_src"] == 1) | (cls["chir_dst"] == 1)],
                                  re[(cls["chir_src"] == 1) | (cls["chir_dst"] == 1)])
    ctx_e.close()
    ctx_m.close()
