"""The allocation-free MSD reduction (kmc_reduce_samples_noalloc, a static device scratch per
GPU) against kmc_reduce_samples, bit for bit, through the C ABI: two calls in a row, the second
with a smaller input (fewer excitons, bins and batches) so that stale partials of the first
would show; then the capacity limit, refused on the host."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-noalloc]", *a, file=sys.stderr, flush=True)


def _run(kmc, nw, E, bins, hops, seed):
    w = nw.film_small()
    net = nw.build_network(w)
    th, z, a1, a2, rates = nw.forster_tables(1)
    ctx = nw.make_context(w, net, seed=seed)
    ctx.set_rate_table(th, z, a1, a2, rates)
    ctx.build_cells()
    ctx.build_rates(kmc.RATE_TABLE)
    ctx.inject(E)
    t0 = float(np.median(ctx.get_excitons()["t_next"]))
    ctx.set_samples(bins, t0 * hops / 3)
    ctx.run(hops)
    return ctx


def test_noalloc_reduce_equals_reduce(kmc, nw):
    say("noalloc reduce: film_5k_sites, 2000 x 3000 hops, 64 bins, 32 batches")
    a = _run(kmc, nw, 2000, 64, 3000, 31)
    ref, _ = a.reduce_samples(32)
    got, ms = a.reduce_samples_noalloc(32)
    assert np.array_equal(got, ref) and ms >= 0.0
    assert np.all(ref[:, :, 0].sum(1) > 0)
    say("noalloc reduce: second call, 300 x 1000 hops, 8 bins, 4 batches")
    b = _run(kmc, nw, 300, 8, 1000, 32)
    ref_b, _ = b.reduce_samples(4)
    got_b, _ = b.reduce_samples_noalloc(4)
    assert np.array_equal(got_b, ref_b)
    # the first context again, after the smaller call
    assert np.array_equal(a.reduce_samples_noalloc(32)[0], ref)
    # 64 bins x 103 batches x 5 > 32768 doubles: refused before any device work
    with pytest.raises(kmc.KmcError) as e:
        a.reduce_samples_noalloc(103)
    assert e.value.code == kmc.ERR_CAPACITY
    assert np.array_equal(a.reduce_samples_noalloc(102)[0], a.reduce_samples(102)[0])
    a.close()
    b.close()
