"""Learned surrogate of the first-order transfer rate (kmc_mlp in include/cntkmc.h).

No reference counterpart: it stands beside exciton_transfer::first_order
(src/exciton_transfer/exciton_transfer.cpp:380-441), whose GPU restatement (exciton.Transfer)
labels the training pairs and recomputes every pair the surrogate flags.

Pipeline (train_surrogate):
  1. pairs: a mixture of segment pairs within a cutoff sphere (isotropic and in-plane
     orientations, mapped to (theta, z, a1, a2) by the reference's pair geometry,
     scatterer.cpp:44-61), collinear same-tube pairs and pairs uniform over the clamp box;
     chirality classes and temperatures drawn per group;
  2. labels: log10 of the exact rate (clamped to the box like the table), floored at
     LOG10_FLOOR (rates below 1 /s are ~12 orders below any site's total rate);
  3. training on the GPU by the library's BF16 MFMA forward/backward kernels with Adam;
  4. a held-out set gives the error against the FP64 rates and the residual map: bins of
     (pair class, T, z, theta, max|a|) in which more than 10% of the held-out pairs exceeded
     the tolerance (held-out p90 residual > tol), or that held fewer than 4 held-out pairs,
     send their pairs to the exact routine.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import exciton, kmc

CHIRALITIES = [(7, 5), (8, 6), (9, 7)]   # surrogate classes 0, 1, 2
IN, WIDTH, HIDDEN, N_PARAMS = 16, 128, 4, 51841
N_FEATURES = 13   # feature slots in use (the rest are zero)
FLAG_DOMAIN, FLAG_RESIDUAL = 1, 2
LOG10_FLOOR = 0.0
DEFAULT_MODEL = Path(__file__).resolve().parent / "data" / "rate_surrogate.npz"


class _Domain(C.Structure):
    _fields_ = [("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3), ("t_lo", C.c_double),
                ("t_hi", C.c_double), ("y_shift", C.c_double), ("y_scale", C.c_double)]


@dataclass
class Domain:
    box_lo: tuple = (1.5e-9, -10e-9, -10e-9)   # input.json:161-163 table box: z, shift 1, shift 2
    box_hi: tuple = (10e-9, 10e-9, 10e-9)
    t_lo: float = 150.0
    t_hi: float = 400.0
    y_shift: float = 7.0
    y_scale: float = 4.0

    def c(self):
        d = _Domain()
        for k in range(3):
            d.box_lo[k] = self.box_lo[k]
            d.box_hi[k] = self.box_hi[k]
        d.t_lo, d.t_hi, d.y_shift, d.y_scale = self.t_lo, self.t_hi, self.y_shift, self.y_scale
        return d


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)
_SIGS = {
    "kmc_mlp_create": (C.c_int, [C.c_int32, C.POINTER(_Domain), C.POINTER(_P)]),
    "kmc_mlp_destroy": (C.c_int, [_P]),
    "kmc_mlp_last_error": (C.c_char_p, []),
    "kmc_mlp_set_params": (C.c_int, [_P, _F]),
    "kmc_mlp_get_params": (C.c_int, [_P, _F]),
    "kmc_mlp_set_residual_map": (C.c_int, [_P, _I32, _U8]),
    "kmc_mlp_features": (C.c_int, [_P, C.c_int64, _D, _D, _D, _D, _I32, _I32, _D, _F]),
    "kmc_mlp_forward": (C.c_int, [_P, C.c_int64, _F, _F]),
    "kmc_mlp_loss_grad": (C.c_int, [_P, C.c_int64, _F, _F, _D, _F]),
    "kmc_mlp_adam_step": (C.c_int, [_P, _F, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_int64]),
    "kmc_mlp_train": (C.c_int, [_P, C.c_int64, _F, _F, C.c_int32, C.c_int32, C.c_double,
                                C.c_double, C.c_uint64, _D, _D]),
    "kmc_mlp_rates_device": (C.c_int, [_P, C.c_int64, _P, _P, _P, _P, C.c_int32, C.c_int32,
                                       C.c_double, _P, _P]),
    "kmc_mlp_rates": (C.c_int, [_P, C.c_int64, _D, _D, _D, _D, C.c_int32, C.c_int32, C.c_double,
                                _D, _U8, _D]),
    "kmc_mlp_table": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_double, _P, C.c_int32, _D,
                                C.c_int32, _D, C.c_int32, _D, C.c_int32, _D, _D, _I64, _D]),
}
_bound = False


def library():
    global _bound
    lib = kmc.library()
    if not _bound:
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return lib


def exported_symbols():
    return list(_SIGS)


def _p(a, ct=_D):
    return None if a is None else a.ctypes.data_as(ct)


def _f64(v):
    return np.ascontiguousarray(v, np.float64).ravel()


class RateSurrogate:
    """The device MLP: parameters, training, fused inference and surrogate-built tables."""

    def __init__(self, domain: Domain | None = None, device: int = 0):
        self.domain = domain or Domain()
        self.lib = library()
        h = _P()
        self._ck(self.lib.kmc_mlp_create(device, C.byref(self.domain.c()), C.byref(h)))
        self.h = h
        self.rmap_dims = None
        self.rmap = None
        self.meta = {}

    def _ck(self, rc):
        if rc < 0:
            raise kmc.KmcError(rc, self.lib.kmc_mlp_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.kmc_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters -----------------------------------------------------------------------
    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == N_PARAMS
        self._ck(self.lib.kmc_mlp_set_params(self.h, _p(p, _F)))

    def get_params(self):
        p = np.zeros(N_PARAMS, np.float32)
        self._ck(self.lib.kmc_mlp_get_params(self.h, _p(p, _F)))
        return p

    def set_residual_map(self, dims, flags):
        if flags is None:
            self._ck(self.lib.kmc_mlp_set_residual_map(self.h, None, None))
            self.rmap_dims = self.rmap = None
            return
        d = np.ascontiguousarray(dims, np.int32)
        f = np.ascontiguousarray(flags, np.uint8).ravel()
        assert f.size == int(np.prod(d))
        self._ck(self.lib.kmc_mlp_set_residual_map(self.h, _p(d, _I32), _p(f, _U8)))
        self.rmap_dims, self.rmap = d, f

    # -- training path --------------------------------------------------------------------
    def features(self, theta, z, a1, a2, cd, ca, T):
        g = [_f64(v) for v in (theta, z, a1, a2)]
        n = g[0].size
        cd = np.ascontiguousarray(np.broadcast_to(cd, n), np.int32)
        ca = np.ascontiguousarray(np.broadcast_to(ca, n), np.int32)
        T = np.ascontiguousarray(np.broadcast_to(T, n), np.float64)
        f = np.zeros((n, IN), np.float32)
        self._ck(self.lib.kmc_mlp_features(self.h, n, *[_p(v) for v in g], _p(cd, _I32),
                                           _p(ca, _I32), _p(T), _p(f, _F)))
        return f

    def forward(self, feat):
        feat = np.ascontiguousarray(feat, np.float32)
        y = np.zeros(feat.shape[0], np.float32)
        self._ck(self.lib.kmc_mlp_forward(self.h, feat.shape[0], _p(feat, _F), _p(y, _F)))
        return y

    def loss_grad(self, feat, target):
        feat = np.ascontiguousarray(feat, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_double(0.0)
        g = np.zeros(N_PARAMS, np.float32)
        self._ck(self.lib.kmc_mlp_loss_grad(self.h, feat.shape[0], _p(feat, _F), _p(t, _F),
                                            C.byref(loss), _p(g, _F)))
        return loss.value, g

    def adam_step(self, grad, lr, beta1=0.9, beta2=0.999, eps=1e-8, step=1):
        g = np.ascontiguousarray(grad, np.float32)
        self._ck(self.lib.kmc_mlp_adam_step(self.h, _p(g, _F), lr, beta1, beta2, eps, step))

    def train(self, feat, target, epochs, batch=8192, lr=2e-3, lr_min=2e-5, seed=1):
        feat = np.ascontiguousarray(feat, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        losses = np.zeros(epochs)
        ms = C.c_double(0.0)
        self._ck(self.lib.kmc_mlp_train(self.h, feat.shape[0], _p(feat, _F), _p(t, _F), epochs,
                                        batch, lr, lr_min, seed, _p(losses), C.byref(ms)))
        return losses, ms.value

    # -- inference ------------------------------------------------------------------------
    def rates(self, theta, z, a1, a2, cls_d, cls_a, T):
        """(rate, flag, kernel_ms) for one chirality pair and temperature."""
        g = [_f64(v) for v in (theta, z, a1, a2)]
        n = g[0].size
        out = np.zeros(n)
        fl = np.zeros(n, np.uint8)
        ms = C.c_double(0.0)
        self._ck(self.lib.kmc_mlp_rates(self.h, n, *[_p(v) for v in g], cls_d, cls_a, float(T),
                                        _p(out), _p(fl, _U8), C.byref(ms)))
        return out, fl, ms.value

    def table(self, cls_d, cls_a, T, fallback, theta, z, a1, a2):
        """(table, nodes recomputed exactly, wall ms) on the grid."""
        g = [_f64(v) for v in (theta, z, a1, a2)]
        out = np.zeros([v.size for v in g])
        nfb = C.c_int64(0)
        ms = C.c_double(0.0)
        self._ck(self.lib.kmc_mlp_table(self.h, cls_d, cls_a, float(T),
                                        None if fallback is None else fallback.h,
                                        g[0].size, _p(g[0]), g[1].size, _p(g[1]), g[2].size,
                                        _p(g[2]), g[3].size, _p(g[3]), _p(out), C.byref(nfb),
                                        C.byref(ms)))
        return out, nfb.value, ms.value

    def log10_of(self, y):
        return self.domain.y_shift + self.domain.y_scale * np.asarray(y, np.float64)

    # -- files ----------------------------------------------------------------------------
    def save(self, path):
        d = self.domain
        np.savez_compressed(
            path, params=self.get_params(),
            domain=np.array(list(d.box_lo) + list(d.box_hi) + [d.t_lo, d.t_hi, d.y_shift, d.y_scale]),
            rmap_dims=np.zeros(5, np.int32) if self.rmap_dims is None else self.rmap_dims,
            rmap=np.zeros(0, np.uint8) if self.rmap is None else self.rmap,
            meta=np.frombuffer(json.dumps(self.meta).encode(), np.uint8))

    @classmethod
    def load(cls, path=DEFAULT_MODEL, device: int = 0):
        z = np.load(path, allow_pickle=False)
        v = z["domain"]
        dom = Domain(tuple(v[0:3]), tuple(v[3:6]), float(v[6]), float(v[7]), float(v[8]), float(v[9]))
        s = cls(dom, device)
        s.set_params(z["params"])
        if z["rmap"].size:
            s.set_residual_map(z["rmap_dims"], z["rmap"])
        s.meta = json.loads(bytes(z["meta"]).decode()) if z["meta"].size else {}
        return s


def film_tables(chiralities, temperature: float, surrogate: RateSurrogate | None = None,
                device: int = 0, length_cells: int = 10):
    """The n^2 donor/acceptor rate tables of a film's chiralities on the default grid
    (create_davoody_scatt_table, src/monte_carlo/monte_carlo.cpp:64-154), built by the
    surrogate with the exact first-order rate recomputing the flagged nodes.  Returns
    (theta, z, a1, a2, rates [n^2][n_theta][n_z][n_a1][n_a2], nodes recomputed exactly)."""
    sur = surrogate or RateSurrogate.load(device=device)
    th, z, a1, a2 = kmc.default_grids()
    models = {c: exciton.CntModel(*c, length_cells, device=device) for c in chiralities}
    tabs, n_exact = [], 0
    try:
        for ci in chiralities:
            for cj in chiralities:
                tr = exciton.Transfer(models[ci], models[cj], temperature, device=device)
                try:
                    tr.set_clamp(sur.domain.box_lo, sur.domain.box_hi)
                    tab, nfb, _ = sur.table(CHIRALITIES.index(ci), CHIRALITIES.index(cj),
                                            temperature, tr, th, z, a1, a2)
                finally:
                    tr.close()
                tabs.append(tab)
                n_exact += nfb
    finally:
        for m in models.values():
            m.close()
        if surrogate is None:
            sur.close()
    return th, z, a1, a2, np.stack(tabs), n_exact


# ---------------------------------------------------------------------------------------------
# training data
# ---------------------------------------------------------------------------------------------
def pair_geometry(pi, oi, pj, oj):
    """(theta, z, a1, a2) of segment pairs by the reference's formulas (scatterer.cpp:44-61)."""
    d = pi - pj
    c = np.clip(np.einsum("ij,ij->i", oi, oj), -1.0, 1.0)
    par = c == 1.0
    s = np.where(par, 1.0, 1.0 - c * c)
    y1 = np.einsum("ij,ij->i", oi, d)
    y2 = np.einsum("ij,ij->i", oj, d)
    a1 = np.where(par, 0.0, (y1 + y2 * c) / s)
    a2 = np.where(par, y1, (y2 + y1 * c) / s)
    zv = (a1[:, None] * oi + pi) - (a2[:, None] * oj + pj)
    zpar = d - y1[:, None] * oi
    z = np.where(par, np.linalg.norm(zpar, axis=1), np.linalg.norm(zv, axis=1))
    th = np.where(par, 0.0, np.arccos(c))
    return th, z, a1, a2


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def sample_geometries(rng, n, domain: Domain, cutoff=20e-9, min_distance=0.4e-9):
    """Pair geometries, clamped to the box, from a mixture that covers where film pairs lie:
    35% isotropic segment pairs within the cutoff sphere, 30% pairs with in-plane orientations
    (the films' tubes lie in the x-y plane), 5% collinear same-tube pairs (theta = 0, a1 = 0)
    and 30% uniform over the box.  Cutoff-sphere pairs are mapped to (theta, z, a1, a2) by the
    reference's pair geometry (scatterer.cpp:44-61)."""
    n_iso, n_pl, n_col = int(0.35 * n), int(0.30 * n), int(0.05 * n)
    n_box = n - n_iso - n_pl - n_col
    parts = []
    for k, planar in ((n_iso, False), (n_pl, True)):
        r = np.maximum(cutoff * rng.uniform(0, 1, k) ** (1 / 3), min_distance)
        d = _unit(rng, k) * r[:, None]
        if planar:
            a, b = rng.uniform(0, 2 * math.pi, k), rng.uniform(0, 2 * math.pi, k)
            oi = np.stack([np.cos(a), np.sin(a), np.zeros(k)], 1)
            oj = np.stack([np.cos(b), np.sin(b), np.zeros(k)], 1)
        else:
            oi, oj = _unit(rng, k), _unit(rng, k)
        parts.append(pair_geometry(d, oi, np.zeros((k, 3)), oj))
    parts.append((np.zeros(n_col), np.zeros(n_col), np.zeros(n_col),
                  rng.uniform(-cutoff, cutoff, n_col)))
    lo, hi = domain.box_lo, domain.box_hi
    parts.append((rng.uniform(0, math.pi, n_box), rng.uniform(lo[0], hi[0], n_box),
                  rng.uniform(lo[1], hi[1], n_box), rng.uniform(lo[2], hi[2], n_box)))
    g = [np.concatenate([p[q] for p in parts]) for q in range(4)]
    # clamp like the exact routine, so every stored geometry is the one that was labelled
    for q in range(3):
        g[q + 1] = np.clip(g[q + 1], lo[q], hi[q])
    perm = rng.permutation(n)
    return [v[perm] for v in g]


def make_dataset(models, n, seed, domain: Domain, temps_per_class=96, log=print):
    """Exactly labelled pairs: dict(theta, z, a1, a2, cd, ca, T, log10_rate, seconds)."""
    rng = np.random.default_rng(seed)
    nc = len(CHIRALITIES)
    groups = [(cd, ca, float(T)) for cd in range(nc) for ca in range(nc)
              for T in rng.uniform(domain.t_lo, domain.t_hi, temps_per_class)]
    per = int(math.ceil(n / len(groups)))
    cols = {k: [] for k in ("theta", "z", "a1", "a2", "cd", "ca", "T", "log10_rate")}
    t0 = time.time()
    gpu_ms = 0.0
    for gi, (cd, ca, T) in enumerate(groups):
        th, z, a1, a2 = sample_geometries(rng, per, domain)
        tr = exciton.Transfer(models[CHIRALITIES[cd]], models[CHIRALITIES[ca]], T)
        tr.set_clamp(domain.box_lo, domain.box_hi)
        r = tr.rates(th, z, a1, a2)
        gpu_ms += tr.last_ms
        tr.close()
        lab = np.log10(np.maximum(r, 10.0 ** LOG10_FLOOR))
        for k, v in (("theta", th), ("z", z), ("a1", a1), ("a2", a2), ("log10_rate", lab)):
            cols[k].append(v)
        cols["cd"].append(np.full(per, cd, np.int32))
        cols["ca"].append(np.full(per, ca, np.int32))
        cols["T"].append(np.full(per, T))
        if log and (gi + 1) % 384 == 0:
            log(f"labels: {gi + 1}/{len(groups)} groups, {time.time() - t0:.1f} s")
    out = {k: np.concatenate(v)[:n] for k, v in cols.items()}
    out["seconds"] = time.time() - t0
    out["exact_kernel_seconds"] = gpu_ms / 1e3
    return out


def init_params(seed):
    """Uniform He initialisation of the hidden layers, small output layer, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    off = 0
    for l in range(HIDDEN):
        fan = IN if l == 0 else WIDTH
        used = N_FEATURES if l == 0 else WIDTH
        lim = math.sqrt(6.0 / used)
        w = rng.uniform(-lim, lim, (fan, WIDTH)).astype(np.float32)
        if l == 0:
            w[N_FEATURES:] = 0.0
        p[off:off + fan * WIDTH] = w.ravel()
        off += fan * WIDTH + WIDTH
    p[off:off + WIDTH] = rng.uniform(-1, 1, WIDTH) * math.sqrt(3.0 / WIDTH)
    return p


def relative_error(log10_pred, log10_exact):
    """|rate_pred / rate_exact - 1| with both floored at LOG10_FLOOR."""
    a = np.maximum(np.asarray(log10_pred, np.float64), LOG10_FLOOR)
    b = np.maximum(np.asarray(log10_exact, np.float64), LOG10_FLOOR)
    return np.abs(np.expm1((a - b) * math.log(10.0)))


def bin_indices(dims, cd, ca, T, theta, z, a1, a2, domain: Domain):
    """Flattened residual-map bin of each pair (the binning kmc_mlp_rates applies)."""
    lo, hi = domain.box_lo, domain.box_hi

    def b(u, n):
        return np.clip((u * n).astype(np.int64), 0, n - 1)

    zc = np.clip(z, lo[0], hi[0])
    c1 = np.abs(np.clip(a1, lo[1], hi[1]))
    c2 = np.abs(np.clip(a2, lo[2], hi[2]))
    amax = max(abs(lo[1]), abs(hi[1]), abs(lo[2]), abs(hi[2]))
    iT = b((np.asarray(T) - domain.t_lo) / (domain.t_hi - domain.t_lo), dims[1])
    iz = b((zc - lo[0]) / (hi[0] - lo[0]), dims[2])
    ith = b(np.asarray(theta) / math.pi, dims[3])
    ia = b((np.maximum(c1, c2) / amax) ** 2, dims[4])
    pc = np.asarray(cd) * 3 + np.asarray(ca)
    return (((pc * dims[1] + iT) * dims[2] + iz) * dims[3] + ith) * dims[4] + ia


def residual_map(dims, bins, rel_err, tol=0.05, quantile=0.9, min_count=4):
    """Flag the bins whose held-out relative error at `quantile` exceeds tol (more than
    1 - quantile of their held-out pairs beyond tol), and bins with fewer than min_count
    held-out pairs, whose error is unknown."""
    size = int(np.prod(dims))
    cnt = np.bincount(bins, minlength=size)
    over = np.bincount(bins, weights=(np.asarray(rel_err) > tol).astype(np.float64), minlength=size)
    return ((over > (1.0 - quantile) * cnt) | (cnt < min_count)).astype(np.uint8)


def validate(s, models, n_val, seed, tol=0.05, rmap_dims=(9, 10, 8, 8, 4), temps_per_class=128,
             log=print):
    """Fresh exactly-labelled held-out pairs: the error of the surrogate against the FP64
    rates and the residual map built from it (installed on s).  Returns a stats dict."""
    dom = s.domain
    va = make_dataset(models, n_val, seed, dom, temps_per_class=temps_per_class, log=log)
    f_va = s.features(va["theta"], va["z"], va["a1"], va["a2"], va["cd"], va["ca"], va["T"])
    err = relative_error(s.log10_of(s.forward(f_va)), va["log10_rate"])
    bins = bin_indices(rmap_dims, va["cd"], va["ca"], va["T"], va["theta"], va["z"], va["a1"],
                       va["a2"], dom)
    flags = residual_map(rmap_dims, bins, err, tol)
    s.set_residual_map(rmap_dims, flags)
    vf = flags[bins] != 0
    per = {}
    for cd in range(3):
        for ca in range(3):
            m = (va["cd"] == cd) & (va["ca"] == ca)
            per[f"{CHIRALITIES[cd]}->{CHIRALITIES[ca]}"] = {
                "median": float(np.median(err[m])), "p99": float(np.quantile(err[m], 0.99))}
    return {"n_val": int(n_val), "val_median_rel": float(np.median(err)),
            "val_p95_rel": float(np.quantile(err, 0.95)), "val_p99_rel": float(np.quantile(err, 0.99)),
            "val_frac_over_tol": float(np.mean(err > tol)), "flagged_bins": float(flags.mean()),
            "val_frac_flagged": float(np.mean(vf)),
            "val_frac_over_tol_unflagged": float(np.mean((err > tol) & ~vf)), "per_class": per}


@dataclass
class TrainReport:
    n_train: int
    n_val: int
    label_seconds: float
    exact_kernel_seconds: float
    train_seconds: float
    epochs: int
    final_loss: float
    val_median_rel: float
    val_p95_rel: float
    val_p99_rel: float
    val_frac_over_tol: float
    flagged_bins: float
    val_frac_flagged: float
    val_frac_over_tol_unflagged: float
    per_class: dict = field(default_factory=dict)


def train_chunked(s, feat, target, epochs, batch, lr, lr_min, seed, chunk=100, log=print):
    """Train in chunks of `chunk` epochs: each chunk runs a half cosine between the values the
    global cosine schedule (lr -> lr_min over all epochs) takes at its ends, so the schedule
    follows the global one piecewise while progress is reported between chunks.  Adam's
    moments and step count carry over from chunk to chunk.  Returns (losses, ms)."""
    def sched(e):
        return lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * e / epochs))
    losses, ms = [], 0.0
    for k, e0 in enumerate(range(0, epochs, chunk)):
        e1 = min(epochs, e0 + chunk)
        l, m = s.train(feat, target, e1 - e0, batch, sched(e0), sched(e1), seed + k)
        losses.extend(l.tolist())
        ms += m
        if log:
            log(f"epochs {e1}/{epochs}: loss {l[-1]:.3e} ({ms / 1e3:.1f} s)")
    return np.array(losses), ms


def train_surrogate(n_train=2_000_000, n_val=2_000_000, epochs=1200, batch=8192, lr=3e-3,
                    lr_min=1e-6, seed=1001, tol=0.05, rmap_dims=(9, 10, 8, 8, 4), device=0,
                    models=None, log=print):
    """Label, train, validate, build the residual map.  Returns (RateSurrogate, TrainReport)."""
    dom = Domain()
    models = models or {c: exciton.CntModel(*c, 10) for c in CHIRALITIES}
    tr = make_dataset(models, n_train, seed, dom, temps_per_class=384, log=log)
    lab = tr["log10_rate"]
    dom.y_shift = float(np.mean(lab))
    dom.y_scale = float(np.std(lab))
    s = RateSurrogate(dom, device)
    s.set_params(init_params(seed))
    f_tr = s.features(tr["theta"], tr["z"], tr["a1"], tr["a2"], tr["cd"], tr["ca"], tr["T"])
    y_tr = ((lab - dom.y_shift) / dom.y_scale).astype(np.float32)
    losses, ms = train_chunked(s, f_tr, y_tr, epochs, batch, lr, lr_min, seed, log=log)
    log(f"trained {epochs} epochs in {ms / 1e3:.1f} s, loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    st = validate(s, models, n_val, seed + 7919, tol, rmap_dims, log=log)
    rep = TrainReport(n_train, n_val, tr["seconds"], tr["exact_kernel_seconds"], ms / 1e3, epochs,
                      float(losses[-1]), st["val_median_rel"], st["val_p95_rel"], st["val_p99_rel"],
                      st["val_frac_over_tol"], st["flagged_bins"], st["val_frac_flagged"],
                      st["val_frac_over_tol_unflagged"], st["per_class"])
    s.meta = {"train": rep.__dict__, "seed": seed, "tol": tol,
              "chiralities": CHIRALITIES, "loss_curve": [float(v) for v in losses]}
    return s, rep
