"""GPU tests of the rate-surrogate kernels against the CPU FP64 oracle MLP (oracle/mlp_oracle.py).

The oracle rounds to BF16 at the same points as the device; what remains is FP32 accumulation
order and the hardware exp2/reciprocal inside SiLU, which can move a BF16 rounding by one unit
now and then.  Tolerances: network output |dy| <= 2e-3 against the oracle and <= 5e-3 between
the fused inference kernel and the training forward (outputs are O(1); the two device paths
sum in different orders, so their roundings flip independently); gradients per parameter block
within 0.2% in L2 norm; Adam 1e-5 relative.  A transposed operand, a wrong bias or a dropped
layer gives O(1) differences.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-surrogate]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def S():
    from cnt_film_monte_carlo_amd import surrogate
    return surrogate


@pytest.fixture(scope="module")
def MO():
    from oracle import mlp_oracle
    return mlp_oracle


@pytest.fixture(scope="module")
def net(S):
    rng = np.random.default_rng(11)
    p = S.init_params(5)
    # nonzero biases everywhere so that every bias path is exercised
    for name, off, shape in __import__("oracle.mlp_oracle", fromlist=["x"]).LAYOUT:
        if name.startswith("b"):
            p[off:off + int(np.prod(shape))] = rng.normal(0, 0.3, shape)
    s = S.RateSurrogate(S.Domain(y_shift=10.0, y_scale=2.0))
    s.set_params(p)
    return s, p


def geometries(rng, n, S):
    d = S.Domain()
    return (rng.uniform(0, math.pi, n), rng.uniform(0.5e-9, 12e-9, n), rng.uniform(-12e-9, 12e-9, n),
            rng.uniform(-12e-9, 12e-9, n))


def test_features_match_oracle(S, MO, net):
    s, _ = net
    rng = np.random.default_rng(1)
    th, z, a1, a2 = geometries(rng, 2000, S)
    cd, ca = rng.integers(0, 3, 2000), rng.integers(0, 3, 2000)
    T = rng.uniform(150, 400, 2000)
    f = s.features(th, z, a1, a2, cd, ca, T)
    d = s.domain
    g = MO.features(th, z, a1, a2, cd, ca, T, d.box_lo, d.box_hi, d.t_lo, d.t_hi)
    np.testing.assert_allclose(f, g, rtol=0, atol=2e-7)
    assert np.all(f[:, 13:] == 0) and np.all(np.abs(f[:, 2:5]) <= 1.0)
    assert np.all(np.abs(f[:, 11:13]) <= 1.0 + 1e-6)


def test_forward_matches_oracle(S, MO, net):
    s, p = net
    rng = np.random.default_rng(2)
    n = 1000  # not a multiple of the 128-row batch granule
    th, z, a1, a2 = geometries(rng, n, S)
    f = s.features(th, z, a1, a2, rng.integers(0, 3, n), rng.integers(0, 3, n), rng.uniform(150, 400, n))
    y = s.forward(f)
    yo, _ = MO.forward(p.astype(np.float64), f.astype(np.float64), bf=True)
    err = np.abs(y - yo)
    say(f"forward: max |dy| {err.max():.2e}, median {np.median(err):.2e}, |y| ~ {np.abs(yo).mean():.2f}")
    assert err.max() <= 2e-3 and np.median(err) <= 2e-4
    # the plain FP64 network is close too (BF16 operands cost about 1e-2)
    yf, _ = MO.forward(p.astype(np.float64), f.astype(np.float64), bf=False)
    assert np.abs(y - yf).max() < 0.1


def test_inference_kernel_matches_training_forward(S, net):
    s, _ = net
    rng = np.random.default_rng(3)
    n = 5000
    th, z, a1, a2 = geometries(rng, n, S)
    for cd, ca, T in ((0, 0, 300.0), (1, 2, 170.0), (2, 1, 390.0)):
        rate, flag, ms = s.rates(th, z, a1, a2, cd, ca, T)
        assert np.all(flag == 0) and np.all(rate > 0)
        y = s.forward(s.features(th, z, a1, a2, cd, ca, T))
        d = np.abs(np.log10(rate) - s.log10_of(y)) / s.domain.y_scale
        say(f"fused vs training forward ({cd},{ca},{T}): max |dy| {d.max():.2e}")
        assert d.max() <= 5e-3 and np.median(d) <= 2e-4


def test_loss_grad_matches_oracle(S, MO, net):
    s, p = net
    rng = np.random.default_rng(4)
    for n in (512, 300):
        th, z, a1, a2 = geometries(rng, n, S)
        f = s.features(th, z, a1, a2, rng.integers(0, 3, n), rng.integers(0, 3, n),
                       rng.uniform(150, 400, n))
        t = rng.normal(0, 1, n).astype(np.float32)
        loss, g = s.loss_grad(f, t)
        lo, go = MO.loss_grad(p.astype(np.float64), f.astype(np.float64), t.astype(np.float64), bf=True)
        assert loss == pytest.approx(lo, rel=2e-3)
        for name, off, shape in MO.LAYOUT:
            k = int(np.prod(shape))
            a, b = g[off:off + k].astype(np.float64), go[off:off + k]
            if name == "W0":  # feature slots 13..15 are always zero: their rows get no gradient
                a, b = a.reshape(shape)[:13], b.reshape(shape)[:13]
                assert np.all(g[off:off + k].reshape(shape)[13:] == 0)
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            say(f"n={n} grad {name}: rel L2 {rel:.2e}")
            assert rel < 2e-3, name


def test_adam_matches_oracle(S, MO, net):
    s, p = net
    rng = np.random.default_rng(5)
    s.set_params(p)
    m = np.zeros(p.size)
    v = np.zeros(p.size)
    q = p.astype(np.float64)
    for step in (1, 2, 3):
        g = rng.normal(0, 1e-2, p.size).astype(np.float32)
        s.adam_step(g, 1e-3, 0.9, 0.999, 1e-8, step)
        q, m, v = MO.adam(q, g.astype(np.float64), m, v, step, 1e-3)
    np.testing.assert_allclose(s.get_params(), q, rtol=1e-5, atol=1e-7)
    s.set_params(p)


def test_training_fits_a_teacher(S, MO):
    # targets from a random FP64 teacher network: the device training must drive the loss down
    rng = np.random.default_rng(6)
    n = 65536
    s = S.RateSurrogate(S.Domain(y_shift=0.0, y_scale=1.0))
    th, z, a1, a2 = geometries(rng, n, S)
    f = s.features(th, z, a1, a2, rng.integers(0, 3, n), rng.integers(0, 3, n), rng.uniform(150, 400, n))
    teacher = MO.init_params(99)
    t, _ = MO.forward(teacher, f.astype(np.float64), bf=False)
    t = ((t - t.mean()) / t.std()).astype(np.float32)
    s.set_params(S.init_params(7))
    l0, _ = s.loss_grad(f[:4096], t[:4096])
    losses, ms = s.train(f, t, epochs=30, batch=1024, lr=2e-3, lr_min=1e-4, seed=3)
    l1, _ = s.loss_grad(f[:4096], t[:4096])
    say(f"teacher fit: loss {l0:.3e} -> {l1:.3e} in {ms:.0f} ms ({30 * (n // 1024)} steps)")
    assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]
    assert l1 < l0 / 20


def test_flags_domain_and_residual(S, net):
    s, _ = net
    rng = np.random.default_rng(8)
    n = 3000
    th, z, a1, a2 = geometries(rng, n, S)
    dims = (9, 5, 6, 7, 3)
    T = 280.0
    bins = S.bin_indices(dims, np.zeros(n, int), np.ones(n, int), np.full(n, T), th, z, a1, a2, s.domain)
    target = bins[17]
    flags = np.zeros(int(np.prod(dims)), np.uint8)
    flags[target] = 1
    s.set_residual_map(dims, flags)
    try:
        _, fl, _ = s.rates(th, z, a1, a2, 0, 1, T)
        np.testing.assert_array_equal(fl, np.where(bins == target, S.FLAG_RESIDUAL, 0))
        assert fl[17] == S.FLAG_RESIDUAL
        _, fl, _ = s.rates(th, z, a1, a2, 0, 5, T)       # unknown class
        assert np.all(fl == S.FLAG_DOMAIN)
        _, fl, _ = s.rates(th, z, a1, a2, 0, 1, 420.0)   # outside the trained temperatures
        assert np.all(fl == S.FLAG_DOMAIN)
    finally:
        s.set_residual_map(None, None)


def test_reversing_both_tubes_gives_the_same_rate(S, net):
    # (a1, a2) -> (-a1, -a2) is reversing both tubes; the features are canonicalised, so the
    # two geometries must get bitwise the same rate (both hop directions along a tube included)
    s, _ = net
    rng = np.random.default_rng(9)
    th, z, a1, a2 = geometries(rng, 4000, S)
    a1[:100] = 0.0
    a2[:100] = rng.choice([-10e-9, 10e-9], 100)
    r0, _, _ = s.rates(th, z, a1, a2, 0, 0, 300.0)
    r1, _, _ = s.rates(th, z, -a1, -a2, 0, 0, 300.0)
    np.testing.assert_array_equal(r0, r1)


def test_inference_many_tiles_per_wave(S, MO, net):
    """300,000 pairs are 9,375 tiles of 32: more than one per wave of the one-workgroup-per-CU
    grid, so each wave strides over several tiles.  Every rate equals the one the same kernel
    gives on chunks of 2,000 pairs (one tile per wave), bit for bit, and a sample follows the
    oracle forward to the bar of the fused kernel against the training forward (5e-3, median
    2e-4): the fused kernel sums in its own order."""
    s, p = net
    rng = np.random.default_rng(10)
    n = 300_000
    th, z, a1, a2 = geometries(rng, n, S)
    big, fl, ms = s.rates(th, z, a1, a2, 1, 2, 250.0)
    say(f"many tiles: {n} pairs in {ms:.3f} ms")
    parts = [s.rates(th[i:i + 2000], z[i:i + 2000], a1[i:i + 2000], a2[i:i + 2000], 1, 2, 250.0)
             for i in range(0, n, 2000)]
    np.testing.assert_array_equal(big, np.concatenate([q[0] for q in parts]))
    np.testing.assert_array_equal(fl, np.concatenate([q[1] for q in parts]))
    k = rng.choice(n, 500, replace=False)
    f = s.features(th[k], z[k], a1[k], a2[k], np.ones(500, int), np.full(500, 2), np.full(500, 250.0))
    yo, _ = MO.forward(p.astype(np.float64), f.astype(np.float64), bf=True)
    dy = np.abs(np.log10(big[k]) - s.log10_of(yo)) / s.domain.y_scale
    say(f"  vs oracle: max |dy| {dy.max():.2e}, median {np.median(dy):.2e}")
    assert dy.max() <= 5e-3 and np.median(dy) <= 2e-4
