"""CPU FP64 restatement of the rate-surrogate MLP: features, forward, loss gradient, Adam.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's baseline leg as the
checker of the device kernels; the product never calls it.

The surrogate has no reference counterpart (it stands beside exciton_transfer::first_order,
src/exciton_transfer/exciton_transfer.cpp:380-441).  With `bf16=True` every value the device
feeds to a BF16 matrix core operand is rounded to BF16 (round to nearest even) at the same
point as on the device:
  forward:  x, W0..W3, and each hidden activation h_l = silu(z_l);
  backward: dz_l (the gradient w.r.t. each pre-activation) before it enters either product.
Everything else (biases, the output layer, accumulation) is exact FP64 here and FP32 on the
device.  With `bf16=False` this is the plain FP64 network, pinned by finite differences.
"""
from __future__ import annotations

import math

import numpy as np

IN, WD, NH = 16, 128, 4
CLASSES = 3


def layout():
    """[(name, offset, shape)] of the flat parameter vector."""
    out, off = [], 0
    for l in range(NH):
        fan_in = IN if l == 0 else WD
        out.append((f"W{l}", off, (fan_in, WD)))
        off += fan_in * WD
        out.append((f"b{l}", off, (WD,)))
        off += WD
    out.append(("w4", off, (WD,)))
    off += WD
    out.append(("b4", off, (1,)))
    off += 1
    return out, off


LAYOUT, N_PARAMS = layout()


def unpack(p):
    p = np.asarray(p, np.float64)
    return {name: p[o:o + int(np.prod(s))].reshape(s) for name, o, s in LAYOUT}


def pack(d):
    p = np.zeros(N_PARAMS)
    for name, o, s in LAYOUT:
        p[o:o + int(np.prod(s))] = np.asarray(d[name]).ravel()
    return p


def bf16(x):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(seed=0):
    """He-style uniform init (the device trains from whatever it is given)."""
    rng = np.random.default_rng(seed)
    d = {}
    for l in range(NH):
        fan_in = 13 if l == 0 else WD
        lim = math.sqrt(6.0 / fan_in)
        w = rng.uniform(-lim, lim, (IN if l == 0 else WD, WD))
        if l == 0:
            w[13:] = 0.0
        d[f"W{l}"] = w
        d[f"b{l}"] = np.zeros(WD)
    d["w4"] = rng.uniform(-1, 1, WD) * math.sqrt(3.0 / WD)
    d["b4"] = np.zeros(1)
    return pack(d)


def forward(p, x, bf=True):
    """y[n] and the cache for backward."""
    P = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    h = r(x)
    cache = {"h": [h], "z": []}
    for l in range(NH):
        z = h @ r(P[f"W{l}"]) + P[f"b{l}"]
        h = r(silu(z))
        cache["z"].append(z)
        cache["h"].append(h)
    y = h @ P["w4"] + P["b4"][0]
    return y, cache


def loss_grad(p, x, t, bf=True):
    """Mean squared error and its gradient (flat, the parameter layout)."""
    P = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    y, c = forward(p, x, bf)
    n = len(t)
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / n
    g = {"w4": c["h"][NH].T @ dy, "b4": np.array([dy.sum()])}
    dz = r(dy[:, None] * P["w4"][None, :] * dsilu(c["z"][NH - 1]))
    for l in range(NH - 1, -1, -1):
        g[f"W{l}"] = c["h"][l].T @ dz
        g[f"b{l}"] = dz.sum(0)
        if l > 0:
            dz = r((dz @ r(P[f"W{l}"]).T) * dsilu(c["z"][l - 1]))
    return loss, pack(g)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def features(theta, z, a1, a2, cd, ca, T, box_lo, box_hi, t_lo, t_hi):
    """Feature rows [n][16] (float32), the formulas the device kernel evaluates."""
    theta, z, a1, a2, T = (np.asarray(v, np.float64) for v in (theta, z, a1, a2, T))
    cd, ca = np.asarray(cd), np.asarray(ca)
    n = theta.size
    f = np.zeros((n, IN), np.float64)
    f[:, 0] = np.cos(theta)
    f[:, 1] = np.sin(theta)
    cl = [np.minimum(np.maximum(v, box_lo[k]), box_hi[k]) for k, v in enumerate((z, a1, a2))]
    # reversing both tubes maps (a1, a2) to (-a1, -a2): the sign with a2 > 0 (a1 > 0 on a2 == 0)
    flip = (cl[2] < 0) | ((cl[2] == 0) & (cl[1] < 0))
    cl[1] = np.where(flip, -cl[1], cl[1])
    cl[2] = np.where(flip, -cl[2], cl[2])
    for k, c in enumerate(cl):
        f[:, 2 + k] = (2 * c - (box_lo[k] + box_hi[k])) / (box_hi[k] - box_lo[k])
    for j in range(CLASSES):
        f[:, 5 + j] = (cd == j)
        f[:, 8 + j] = (ca == j)
    f[:, 11] = np.log(T / math.sqrt(t_lo * t_hi)) / math.log(math.sqrt(t_hi / t_lo))
    f[:, 12] = (2.0 / T - (1.0 / t_lo + 1.0 / t_hi)) / (1.0 / t_lo - 1.0 / t_hi)
    return f.astype(np.float32)
