"""GPU tests of ensembles (several networks in one hop launch) and of the schedule autotuner.

An ensemble member's trajectories must be bit-identical to kmc_run on that context alone (same
seed, same exciton ids): sites, visited-site hashes, times, displacements, per-site counts and
MSD records are compared exactly.  The autotuner must leave the state untouched."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-ensemble]", *a, file=sys.stderr, flush=True)


def members(kmc, nw, n=5, seed=77):
    ws = nw.ensemble(n_networks=n, seed=seed, n_min=2_000, n_max=30_000)
    th, z, a1, a2, rates = nw.forster_tables(1)
    out = []
    for i, w in enumerate(ws):
        net = nw.build_network(w)
        for _ in range(2):   # the ensemble's copy and the plain run's copy
            ctx = nw.make_context(w, net, seed=1000 + i)
            ctx.set_rate_table(th, z, a1, a2, rates)
            ctx.build_cells()
            ctx.build_rates(kmc.RATE_TABLE)
            ctx.inject(w.n_excitons)
            ctx.set_samples(16, 2e-5)
            out.append(ctx)
    return ws, out[0::2], out[1::2]


def test_ensemble_equals_plain_runs(kmc, nw):
    ws, ens_ctx, plain = members(kmc, nw)
    say("networks:", [c.n_sites for c in ens_ctx], "excitons:", [w.n_excitons for w in ws])
    ens = kmc.Ensemble(ens_ctx)
    ens.set_hop_schedule(300)
    for H in (700, 1300):          # two calls: resumable like kmc_run
        ens.run(H)
        for c in plain:
            c.run(H)
    for a, b in zip(ens_ctx, plain):
        sa, sb = a.get_excitons(), b.get_excitons()
        for k in ("site", "hops", "hash", "drains"):
            assert np.array_equal(sa[k], sb[k]), k
        assert np.array_equal(sa["t_next"], sb["t_next"])
        assert np.array_equal(sa["disp"], sb["disp"])
        assert np.array_equal(a.get_hop_counts(), b.get_hop_counts())
        ra, rb = a.reduce_samples(4)[0], b.reduce_samples(4)[0]
        assert np.array_equal(ra, rb)
    ens.close()


def test_autotune_leaves_results_unchanged(kmc, nw):
    ws, ens_ctx, plain = members(kmc, nw, n=3, seed=91)
    ens = kmc.Ensemble(ens_ctx)
    best, rates = ens.autotune_schedule(trial_hops=600, candidates=(0, 600, 300, 150))
    say(f"ensemble autotune: best {best}, hops/s {np.round(rates / 1e9, 2)} e9")
    assert best in (0, 600, 300, 150) and np.all(rates > 0)
    ens.run(900)
    best1, rates1 = plain[0].autotune_schedule(trial_hops=600, candidates=(0, 300))
    assert best1 in (0, 300)
    for c in plain:
        c.run(900)
    for a, b in zip(ens_ctx, plain):
        assert np.array_equal(a.get_excitons()["hash"], b.get_excitons()["hash"])
        assert np.array_equal(a.get_hop_counts(), b.get_hop_counts())
    ens.close()


def test_member_changes_are_refused(kmc, nw):
    ws, ens_ctx, _ = members(kmc, nw, n=2, seed=5)
    ens = kmc.Ensemble(ens_ctx)
    ens_ctx[0].inject(ws[0].n_excitons + 1)
    with pytest.raises(kmc.KmcError) as e:
        ens.run(10)
    assert e.value.code == kmc.ERR_STATE
    ens.close()
