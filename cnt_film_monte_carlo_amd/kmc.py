"""ctypes bindings of libcntkmc.so, the C ABI of include/cntkmc.h.

The Python layer is plumbing over the native library: every array goes through the
C ABI unchanged, and every call raises KmcError on a negative status.  The library is
required: importing this module fails loudly when libcntkmc.so has not been built.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcntkmc.so"

OK = 0
ERR_ARG, ERR_STATE, ERR_HIP, ERR_ALLOC, ERR_NODEVICE, ERR_GEOMETRY, ERR_RATE, ERR_CAPACITY = (
    -1, -2, -3, -4, -5, -6, -7, -8)
SITE_SOURCE, SITE_DRAIN = 1, 2
RATE_TABLE, RATE_PAIR = 0, 1
PERIODIC_X, PERIODIC_Y, PERIODIC_Z = 1, 2, 4

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)
_U32 = C.POINTER(C.c_uint32)
_U64 = C.POINTER(C.c_uint64)


class KmcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _Config(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("periodic_mask", C.c_int32),
        ("cutoff", C.c_double),
        ("min_distance", C.c_double),
        ("period_lo", C.c_double * 3),
        ("period_len", C.c_double * 3),
        ("seed", C.c_uint64),
    ]


class _NetParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("n_tubes", C.c_int64),
        ("len_min", C.c_double),
        ("len_max", C.c_double),
        ("box_lo", C.c_double * 3),
        ("box_hi", C.c_double * 3),
        ("site_spacing", C.c_double),
        ("n_chir", C.c_int32),
        ("wrap_mask", C.c_int32),
    ]


_SIGS = {
    "kmc_version": (C.c_char_p, []),
    "kmc_last_error": (C.c_char_p, [_P]),
    "kmc_create": (C.c_int, [C.POINTER(_Config), C.POINTER(_P)]),
    "kmc_destroy": (C.c_int, [_P]),
    "kmc_sync": (C.c_int, [_P]),
    "kmc_set_sites": (C.c_int, [_P, C.c_int64, _D, _D, _D, _D, _D, _D, _I32, _I32, _U8, _U8]),
    "kmc_build_cells": (C.c_int, [_P]),
    "kmc_get_cells": (C.c_int, [_P, _I32, _I64, _I64, _I32]),
    "kmc_set_rate_table": (C.c_int, [_P, C.c_int32, C.c_int32, _D, C.c_int32, _D, C.c_int32, _D,
                                     C.c_int32, _D, _D]),
    "kmc_build_rates": (C.c_int, [_P, C.c_int32]),
    "kmc_get_nnz": (C.c_int, [_P, _I64]),
    "kmc_get_pair_geometry": (C.c_int, [_P, _D, _D, _D, _D, _U8, _U8]),
    "kmc_set_pair_rates": (C.c_int, [_P, _D]),
    "kmc_set_pair_rates_device": (C.c_int, [_P, _P]),
    "kmc_get_csr": (C.c_int, [_P, _I64, _I32, _D, _F, _D]),
    "kmc_set_csr": (C.c_int, [_P, C.c_int64, _I64, _I32, _D, _F]),
    "kmc_inject": (C.c_int, [_P, C.c_int64]),
    "kmc_set_samples": (C.c_int, [_P, C.c_int32, C.c_double]),
    "kmc_set_samples_window": (C.c_int, [_P, C.c_int32, C.c_double, C.c_int32]),
    "kmc_get_sample_window": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "kmc_next_sample_window": (C.c_int, [_P]),
    "kmc_set_trace": (C.c_int, [_P, C.c_int64, C.c_int64]),
    "kmc_run": (C.c_int, [_P, C.c_int64, C.c_double, C.c_int32, _D]),
    "kmc_get_excitons": (C.c_int, [_P, _I32, _D, _U64, _D, _U32, _U64]),
    "kmc_get_hop_counts": (C.c_int, [_P, _U64]),
    "kmc_get_run_stats": (C.c_int, [_P, _I64, _I64]),
    "kmc_get_hop_time": (C.c_int, [_P, _D]),
    "kmc_get_samples": (C.c_int, [_P, _D, _U32, _U32]),
    "kmc_get_trace": (C.c_int, [_P, _I32, _D, _I64]),
    "kmc_reduce_samples": (C.c_int, [_P, C.c_int32, _D, _D]),
    "kmc_reduce_samples_noalloc": (C.c_int, [_P, C.c_int32, _D, _D]),
    "kmc_estimate_diffusion": (C.c_int, [_D, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                         C.c_int32, _D, _D]),
    "kmc_generate_network": (C.c_int, [C.POINTER(_NetParams), C.c_int64, _D, _D, _D, _D, _D, _D,
                                       _I32, _I32, _U8, _I32, _I64]),
    "kmc_forster_table": (C.c_int, [C.c_double, C.c_int32, _D, C.c_int32, _D, C.c_int32, _D,
                                    C.c_int32, _D, _D]),
    "kmc_linspace": (C.c_int, [C.c_double, C.c_double, C.c_int32, _D]),
    "kmc_ff_time": (C.c_int, [C.c_int64, _D, _U32, _D]),
    "kmc_table_rate": (C.c_int, [C.c_int32, C.c_int32, _D, C.c_int32, _D, C.c_int32, _D,
                                 C.c_int32, _D, _D, C.c_int64, _U8, _U8, _D, _D, _D, _D, _D,
                                 _I32]),
    "kmc_save_rate_table": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int64, C.c_int64, C.c_int64,
                                      C.c_int64, _D, _D, _D, _D, _D]),
    "kmc_load_rate_table": (C.c_int, [C.c_char_p, C.c_char_p, _I64, _D, _D, _D, _D, _D]),
    "kmc_read_mesh": (C.c_int, [C.c_char_p, C.c_int64, _I64, _I64, _D, _D, _D, _D, _D, _D, _I32,
                                _I32]),
    "kmc_trim_sites": (C.c_int, [_I64, _D, _D, _D, _D, _D, _D, _I32, _I32, _U8, _U8, _I64, _D, _D,
                                 _D]),
    "kmc_kubo_regions": (C.c_int, [C.c_int64, _D, _D, _D, C.c_int32, _U8, _D, _D]),
    "kmc_build_rates_exact": (C.c_int, [_P, C.c_int32, C.POINTER(_P), _D]),
    "kmc_set_hop_schedule": (C.c_int, [_P, C.c_int64]),
    "kmc_set_velocity": (C.c_int, [_P, C.c_double]),
    "kmc_get_flight": (C.c_int, [_P, _D, _U8, _D]),
    "kmc_contact_init": (C.c_int, [_P, C.c_int32, C.c_int64, C.c_int64, C.c_int64]),
    "kmc_contact_run": (C.c_int, [_P, C.c_double, C.c_int64, C.c_uint64, _I32, _I32, _I64]),
    "kmc_contact_sections": (C.c_int, [_P, _D, _D, _D, _D]),
    "kmc_contact_stats": (C.c_int, [_P, C.POINTER(C.c_uint64), _D]),
    "kmc_contact_particles": (C.c_int, [_P, C.c_int64, _I64, C.POINTER(C.c_uint32), _I32, _D,
                                        C.POINTER(C.c_uint64), _D]),
    "kmc_autotune_schedule": (C.c_int, [_P, C.c_int64, C.c_int32, _I64, _I64, _D]),
    "kmc_ensemble_create": (C.c_int, [C.c_int32, C.POINTER(_P), C.POINTER(_P)]),
    "kmc_ensemble_destroy": (C.c_int, [_P]),
    "kmc_ensemble_last_error": (C.c_char_p, []),
    "kmc_ensemble_set_hop_schedule": (C.c_int, [_P, C.c_int64]),
    "kmc_ensemble_run": (C.c_int, [_P, C.c_int64, C.c_double, C.c_int32, _D]),
    "kmc_ensemble_autotune": (C.c_int, [_P, C.c_int64, C.c_int32, _I64, _I64, _D]),
    "kmc_build_rates_mlp": (C.c_int, [_P, _P, C.c_int32, _I32, C.c_double, C.POINTER(_P), _I64, _D]),
    "kmc_pair_classes": (C.c_int, [_P, C.c_int32, _I64, C.POINTER(_P)]),
    "kmc_get_pair_rates": (C.c_int, [_P, _D]),
}

_lib = None


def library() -> C.CDLL:
    """Load libcntkmc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` (hipcc, gfx950)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _ptr(a: np.ndarray | None, ct):
    if a is None:
        return None
    return a.ctypes.data_as(ct)


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


def _check(code: int, ctx=None):
    if code < 0:
        msg = library().kmc_last_error(ctx).decode()
        raise KmcError(code, msg)
    return code


# ---------------------------------------------------------------------------------------------
@dataclass
class Config:
    cutoff: float
    min_distance: float = 0.4e-9
    periodic_mask: int = 0
    period_lo: tuple = (0.0, 0.0, 0.0)
    period_len: tuple = (0.0, 0.0, 0.0)
    seed: int = 0
    device: int = 0

    def to_c(self) -> _Config:
        c = _Config()
        c.device = self.device
        c.periodic_mask = self.periodic_mask
        c.cutoff = self.cutoff
        c.min_distance = self.min_distance
        for a in range(3):
            c.period_lo[a] = self.period_lo[a]
            c.period_len[a] = self.period_len[a]
        c.seed = self.seed & 0xFFFFFFFFFFFFFFFF
        return c


class Context:
    """One KMC problem on one GPU (wraps kmc_ctx)."""

    def __init__(self, cfg: Config):
        self.lib = library()
        self.cfg = cfg
        h = _P()
        _check(self.lib.kmc_create(C.byref(cfg.to_c()), C.byref(h)))
        self.h = h
        self.n_sites = 0
        self.n_exc = 0
        self.n_bins = 0
        self.t_sample_max = 0.0
        self.n_trace = 0
        self.tr_max = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.kmc_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, code):
        return _check(code, self.h)

    # -- sites / cells / rates ---------------------------------------------------------------
    def set_sites(self, x, y, z, ox, oy, oz, left=None, right=None, chirality=None, flags=None):
        arrs = [_c(v, np.float64) for v in (x, y, z, ox, oy, oz)]
        n = arrs[0].size
        left = None if left is None else _c(left, np.int32)
        right = None if right is None else _c(right, np.int32)
        chir = None if chirality is None else _c(chirality, np.uint8)
        fl = None if flags is None else _c(flags, np.uint8)
        self._ck(self.lib.kmc_set_sites(self.h, n, *[_ptr(a, _D) for a in arrs], _ptr(left, _I32),
                                        _ptr(right, _I32), _ptr(chir, _U8), _ptr(fl, _U8)))
        self.n_sites = n
        # FP32 positions: the hop kernel's displacements are differences of these
        self.site_pos = np.stack(arrs[:3], 1).astype(np.float32)
        self.links = (np.full(n, -1, np.int32) if left is None else left.copy(),
                      np.full(n, -1, np.int32) if right is None else right.copy())

    def build_cells(self):
        self._ck(self.lib.kmc_build_cells(self.h))

    def get_cells(self):
        dims = np.zeros(3, np.int32)
        nc = C.c_int64()
        self._ck(self.lib.kmc_get_cells(self.h, _ptr(dims, _I32), C.byref(nc), None, None))
        start = np.zeros(nc.value + 1, np.int64)
        sites = np.zeros(self.n_sites, np.int32)
        self._ck(self.lib.kmc_get_cells(self.h, None, None, _ptr(start, _I64), _ptr(sites, _I32)))
        return dims, start, sites

    def set_rate_table(self, theta, z_shift, a1, a2, rates):
        g = [_c(v, np.float64) for v in (theta, z_shift, a1, a2)]
        r = _c(rates, np.float64)
        per = g[0].size * g[1].size * g[2].size * g[3].size
        n2 = r.size // per
        n_chir = int(round(math.sqrt(n2)))
        if n_chir * n_chir * per != r.size:
            raise ValueError("rates must hold n_chir^2 tables of the grid size")
        self._ck(self.lib.kmc_set_rate_table(self.h, n_chir, g[0].size, _ptr(g[0], _D), g[1].size,
                                             _ptr(g[1], _D), g[2].size, _ptr(g[2], _D), g[3].size,
                                             _ptr(g[3], _D), _ptr(r, _D)))

    def build_rates(self, mode=RATE_TABLE):
        self._ck(self.lib.kmc_build_rates(self.h, mode))

    @property
    def nnz(self) -> int:
        v = C.c_int64()
        self._ck(self.lib.kmc_get_nnz(self.h, C.byref(v)))
        return v.value

    def get_pair_geometry(self):
        m = self.nnz
        out = [np.zeros(m) for _ in range(4)]
        cs = np.zeros(m, np.uint8)
        cd = np.zeros(m, np.uint8)
        self._ck(self.lib.kmc_get_pair_geometry(self.h, *[_ptr(a, _D) for a in out], _ptr(cs, _U8),
                                                _ptr(cd, _U8)))
        return dict(theta=out[0], z=out[1], a1=out[2], a2=out[3], chir_src=cs, chir_dst=cd)

    def set_pair_rates(self, rates):
        r = _c(rates, np.float64)
        self._ck(self.lib.kmc_set_pair_rates(self.h, _ptr(r, _D)))

    def set_pair_rates_device(self, dev_ptr: int):
        self._ck(self.lib.kmc_set_pair_rates_device(self.h, C.c_void_p(dev_ptr)))

    def build_rates_exact(self, transfers) -> float:
        """Per-pair exact rates; transfers[ci * n_chir + cj] (None -> rate 0).  Returns ms."""
        n2 = len(transfers)
        n_chir = int(round(math.sqrt(n2)))
        arr = (_P * n2)(*[None if t is None else t.h for t in transfers])
        ms = C.c_double(0.0)
        self._ck(self.lib.kmc_build_rates_exact(self.h, n_chir, arr, C.byref(ms)))
        return ms.value

    def build_rates_mlp(self, surrogate, chir_class, temperature: float, transfers):
        """Per-pair surrogate rates with exact fallback for flagged pairs, then the CSR.

        chir_class[i] = surrogate class of the context's chirality id i (-1: unknown);
        transfers[ci * n_chir + cj] recomputes flagged pairs of that class (None allowed where
        no pair is flagged).  Returns (ms, pairs recomputed exactly)."""
        cc = np.ascontiguousarray(chir_class, np.int32)
        n_chir = cc.size
        if len(transfers) != n_chir * n_chir:
            raise ValueError("transfers must have n_chir^2 entries")
        arr = (_P * len(transfers))(*[None if t is None else t.h for t in transfers])
        ms = C.c_double(0.0)
        nfb = C.c_int64(0)
        self._ck(self.lib.kmc_build_rates_mlp(self.h, surrogate.h, n_chir, _ptr(cc, _I32),
                                              float(temperature), arr, C.byref(nfb), C.byref(ms)))
        return ms.value, nfb.value

    def pair_classes(self, n_chir: int):
        start = np.zeros(n_chir * n_chir + 1, np.int64)
        dsel = _P()
        self._ck(self.lib.kmc_pair_classes(self.h, n_chir, _ptr(start, _I64), C.byref(dsel)))
        return start, dsel.value

    def get_pair_rates(self):
        r = np.zeros(self.nnz)
        self._ck(self.lib.kmc_get_pair_rates(self.h, _ptr(r, _D)))
        return r

    def get_csr(self):
        m = self.nnz
        row = np.zeros(self.n_sites + 1, np.int64)
        col = np.zeros(m, np.int32)
        cum = np.zeros(m, np.float64)
        disp = np.zeros((m, 3), np.float32)
        gam = np.zeros(self.n_sites, np.float64)
        self._ck(self.lib.kmc_get_csr(self.h, _ptr(row, _I64), _ptr(col, _I32), _ptr(cum, _D),
                                      _ptr(disp, _F), _ptr(gam, _D)))
        per = np.array([self.cfg.period_len[a] if self.cfg.periodic_mask & (1 << a) else 0.0
                        for a in range(3)])
        return dict(row_ptr=row, col=col, cum=cum, disp=disp, gamma=gam, period=per,
                    site_pos=getattr(self, "site_pos", None))

    def set_csr(self, row_ptr, col, cum, disp):
        row = _c(row_ptr, np.int64)
        self._ck(self.lib.kmc_set_csr(self.h, row.size - 1, _ptr(row, _I64),
                                      _ptr(_c(col, np.int32), _I32), _ptr(_c(cum, np.float64), _D),
                                      _ptr(_c(disp, np.float32), _F)))

    # -- excitons -----------------------------------------------------------------------------
    def inject(self, n_excitons: int):
        self._ck(self.lib.kmc_inject(self.h, int(n_excitons)))
        self.n_exc = int(n_excitons)
        self.n_bins = 0
        self.n_trace = 0

    def set_samples(self, n_bins: int, t_sample_max: float, window: int = 0):
        """MSD sample grid of n_bins times up to t_sample_max; with window > 0 recorded `window`
        bins at a time (records window x excitons instead of n_bins x excitons)."""
        if window:
            self._ck(self.lib.kmc_set_samples_window(self.h, int(n_bins), float(t_sample_max),
                                                     int(window)))
        else:
            self._ck(self.lib.kmc_set_samples(self.h, int(n_bins), float(t_sample_max)))
        self.n_bins = int(n_bins)
        self.t_sample_max = float(t_sample_max)

    def sample_window(self):
        """(first bin, bins) of the open sample window."""
        a, b = C.c_int32(0), C.c_int32(0)
        self._ck(self.lib.kmc_get_sample_window(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def next_sample_window(self):
        self._ck(self.lib.kmc_next_sample_window(self.h))

    def set_trace(self, n_trace: int, max_hops: int):
        self._ck(self.lib.kmc_set_trace(self.h, int(n_trace), int(max_hops)))
        self.n_trace, self.tr_max = int(n_trace), int(max_hops)

    def run(self, hops_per_exciton: int, t_max: float = math.inf, count_sites: bool = True,
            timed: bool = True) -> float:
        ms = C.c_double(0.0)
        self._ck(self.lib.kmc_run(self.h, int(hops_per_exciton), float(t_max), int(count_sites),
                                  C.byref(ms) if timed else None))
        return ms.value

    def run_stats(self, live: bool = True):
        """(hop launches of the last run(), excitons still able to hop after it; None for the
        latter with live=False, which reads nothing from the device)."""
        n, lv = C.c_int64(0), C.c_int64(0)
        self._ck(self.lib.kmc_get_run_stats(self.h, C.byref(n), C.byref(lv) if live else None))
        return n.value, (lv.value if live else None)

    def hop_time(self) -> float:
        """Device ms of the hop-kernel launches alone in the last timed run() (sorts between
        sub-launches left out)."""
        ms = C.c_double(0.0)
        self._ck(self.lib.kmc_get_hop_time(self.h, C.byref(ms)))
        return ms.value

    def set_hop_schedule(self, resort_hops: int):
        """Hops per sub-launch of run() between exciton sorts (0: one launch)."""
        self._ck(self.lib.kmc_set_hop_schedule(self.h, int(resort_hops)))

    def autotune_schedule(self, trial_hops: int = 1250,
                          candidates=(0, 2500, 1250, 625, 312, 156)):
        """Time each schedule on the current state (state restored), install the fastest.
        Returns (best, hops/s per candidate)."""
        c = np.ascontiguousarray(candidates, np.int64)
        rates = np.zeros(c.size)
        best = C.c_int64(0)
        self._ck(self.lib.kmc_autotune_schedule(self.h, int(trial_hops), c.size, _ptr(c, _I64),
                                                C.byref(best), _ptr(rates, _D)))
        return best.value, rates

    def sync(self):
        self._ck(self.lib.kmc_sync(self.h))

    def set_velocity(self, velocity: float):
        """Drift speed along the tubes [m/s] for the excitons of the next inject() (0: hopping
        only)."""
        self._ck(self.lib.kmc_set_velocity(self.h, float(velocity)))

    def get_flight(self):
        """Flight state (v > 0): offset beyond the current site along the heading, heading,
        time of the last hop or injection."""
        E = self.n_exc
        off = np.zeros(E)
        hd = np.zeros(E, np.uint8)
        tl = np.zeros(E)
        self._ck(self.lib.kmc_get_flight(self.h, _ptr(off, _D), _ptr(hd, _U8), _ptr(tl, _D)))
        return dict(offset=off, heading=hd, t_last=tl)

    # ---- contact (steady-state) mode (monte_carlo.h:157-195, 443-684) ----
    def contact_init(self, n_seg: int, c1_pop: int, c2_pop: int, capacity: int):
        """Sections, contact site lists, areas and the initial linear profile; see
        kmc_contact_init."""
        self._ck(self.lib.kmc_contact_init(self.h, int(n_seg), int(c1_pop), int(c2_pop),
                                           int(capacity)))
        self.contact_nseg = int(n_seg)

    def contact_run(self, dt: float, n_steps: int, max_hops_per_step: int = 1 << 20):
        """n_steps steps of dt: (population [n_steps, n_seg], crossings [n_steps, n_seg - 1],
        live particles after each step's repopulation [n_steps])."""
        S = self.contact_nseg
        pop = np.zeros((n_steps, S), np.int32)
        cur = np.zeros((n_steps, S - 1), np.int32)
        live = np.zeros(n_steps, np.int64)
        self._ck(self.lib.kmc_contact_run(self.h, float(dt), int(n_steps), int(max_hops_per_step),
                                          _ptr(pop, _I32), _ptr(cur, _I32), _ptr(live, _I64)))
        return pop, cur, live

    def contact_stats(self):
        """(hops made by all particles, device ms of the step kernels) since contact_init."""
        h = C.c_uint64()
        ms = C.c_double()
        self._ck(self.lib.kmc_contact_stats(self.h, C.byref(h), C.byref(ms)))
        return int(h.value), float(ms.value)

    def contact_sections(self):
        """(area [n_seg], ymin, dy, time)."""
        area = np.zeros(self.contact_nseg)
        v = [C.c_double() for _ in range(3)]
        self._ck(self.lib.kmc_contact_sections(self.h, _ptr(area, _D), *[C.byref(x) for x in v]))
        return area, v[0].value, v[1].value, v[2].value

    def contact_particles(self) -> dict:
        """The live particles in slot order: id, site, t_next, hops, y."""
        n = C.c_int64()
        self._ck(self.lib.kmc_contact_particles(self.h, 0, C.byref(n), None, None, None, None,
                                                None))
        N = n.value
        pid = np.zeros(N, np.uint32)
        site = np.zeros(N, np.int32)
        t = np.zeros(N)
        hops = np.zeros(N, np.uint64)
        y = np.zeros(N)
        self._ck(self.lib.kmc_contact_particles(
            self.h, N, C.byref(n), pid.ctypes.data_as(C.POINTER(C.c_uint32)), _ptr(site, _I32),
            _ptr(t, _D), hops.ctypes.data_as(C.POINTER(C.c_uint64)), _ptr(y, _D)))
        return dict(id=pid, site=site, t_next=t, hops=hops, y=y)

    def get_excitons(self):
        E = self.n_exc
        site = np.zeros(E, np.int32)
        t = np.zeros(E)
        hops = np.zeros(E, np.uint64)
        disp = np.zeros((E, 3))
        dr = np.zeros(E, np.uint32)
        hs = np.zeros(E, np.uint64)
        self._ck(self.lib.kmc_get_excitons(self.h, _ptr(site, _I32), _ptr(t, _D), _ptr(hops, _U64),
                                           _ptr(disp, _D), _ptr(dr, _U32), _ptr(hs, _U64)))
        return dict(site=site, t_next=t, hops=hops, disp=disp, drains=dr, hash=hs)

    def get_hop_counts(self):
        c = np.zeros(self.n_sites, np.uint64)
        self._ck(self.lib.kmc_get_hop_counts(self.h, _ptr(c, _U64)))
        return c

    def get_samples(self):
        T, E = self.sample_window()[1], self.n_exc
        disp = np.zeros((T, E, 3))
        dr = np.zeros((T, E), np.uint32)
        kr = np.zeros(E, np.uint32)
        self._ck(self.lib.kmc_get_samples(self.h, _ptr(disp, _D), _ptr(dr, _U32), _ptr(kr, _U32)))
        return dict(disp=disp, drains=dr, n_recorded=kr)

    def get_trace(self):
        s = np.zeros((self.n_trace, self.tr_max), np.int32)
        t = np.zeros((self.n_trace, self.tr_max))
        n = np.zeros(self.n_trace, np.int64)
        self._ck(self.lib.kmc_get_trace(self.h, _ptr(s, _I32), _ptr(t, _D), _ptr(n, _I64)))
        return dict(sites=s, times=t, n_hops=n)

    def reduce_samples(self, n_batches: int):
        out = np.zeros((self.sample_window()[1], n_batches, 5))
        ms = C.c_double(0.0)
        self._ck(self.lib.kmc_reduce_samples(self.h, int(n_batches), _ptr(out, _D), C.byref(ms)))
        return out, ms.value

    def reduce_samples_noalloc(self, n_batches: int):
        """reduce_samples through the allocation-free entry point (a static device scratch of
        32768 doubles: n_bins * n_batches * 5 must fit)."""
        out = np.zeros((self.sample_window()[1], n_batches, 5))
        ms = C.c_double(0.0)
        self._ck(self.lib.kmc_reduce_samples_noalloc(self.h, int(n_batches), _ptr(out, _D),
                                                     C.byref(ms)))
        return out, ms.value


class Ensemble:
    """Several finished contexts (networks) run by one hop launch (kmc_ensemble)."""

    def __init__(self, contexts):
        self.lib = library()
        self.contexts = list(contexts)       # keep the members alive
        arr = (_P * len(self.contexts))(*[c.h for c in self.contexts])
        h = _P()
        rc = self.lib.kmc_ensemble_create(len(self.contexts), arr, C.byref(h))
        if rc < 0:
            raise KmcError(rc, self.lib.kmc_ensemble_last_error().decode())
        self.h = h

    def _ck(self, rc):
        if rc < 0:
            raise KmcError(rc, self.lib.kmc_ensemble_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.kmc_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_hop_schedule(self, resort_hops: int):
        self._ck(self.lib.kmc_ensemble_set_hop_schedule(self.h, int(resort_hops)))

    def run(self, hops_per_exciton: int, t_max: float = math.inf, count_sites: bool = True) -> float:
        ms = C.c_double(0.0)
        self._ck(self.lib.kmc_ensemble_run(self.h, int(hops_per_exciton), float(t_max),
                                           int(count_sites), C.byref(ms)))
        return ms.value

    def autotune_schedule(self, trial_hops: int = 1250,
                          candidates=(0, 2500, 1250, 625, 312, 156)):
        c = np.ascontiguousarray(candidates, np.int64)
        rates = np.zeros(c.size)
        best = C.c_int64(0)
        self._ck(self.lib.kmc_ensemble_autotune(self.h, int(trial_hops), c.size, _ptr(c, _I64),
                                                C.byref(best), _ptr(rates, _D)))
        return best.value, rates


# ---------------------------------------------------------------------------------------------
# host-only helpers
# ---------------------------------------------------------------------------------------------
@dataclass
class NetworkParams:
    n_tubes: int
    len_min: float
    len_max: float
    box_lo: tuple
    box_hi: tuple
    site_spacing: float = 10e-9
    n_chir: int = 1
    wrap_mask: int = 0
    seed: int = 1001

    def to_c(self) -> _NetParams:
        p = _NetParams()
        p.seed = self.seed
        p.n_tubes = self.n_tubes
        p.len_min = self.len_min
        p.len_max = self.len_max
        for a in range(3):
            p.box_lo[a] = self.box_lo[a]
            p.box_hi[a] = self.box_hi[a]
        p.site_spacing = self.site_spacing
        p.n_chir = self.n_chir
        p.wrap_mask = self.wrap_mask
        return p


def generate_network(params: NetworkParams) -> dict:
    lib = library()
    cp = params.to_c()
    n = C.c_int64()
    _check(lib.kmc_generate_network(C.byref(cp), 0, *([None] * 10), C.byref(n)))
    N = n.value
    out = {k: np.zeros(N) for k in ("x", "y", "z", "ox", "oy", "oz")}
    left = np.zeros(N, np.int32)
    right = np.zeros(N, np.int32)
    chir = np.zeros(N, np.uint8)
    tube = np.zeros(N, np.int32)
    _check(lib.kmc_generate_network(C.byref(cp), N, *[_ptr(out[k], _D) for k in
                                                       ("x", "y", "z", "ox", "oy", "oz")],
                                    _ptr(left, _I32), _ptr(right, _I32), _ptr(chir, _U8),
                                    _ptr(tube, _I32), C.byref(n)))
    out.update(left=left, right=right, chirality=chir, tube=tube, n=N)
    return out


_SITE_KEYS = ("x", "y", "z", "ox", "oy", "oz")


def read_mesh(directory) -> dict:
    """Sites of a CNT mesh directory (six single_cnt.{pos,orient}.{x,y,z}.dat matrices, one row
    per tube, positions in nm) in the generator's dict format, as create_scatterers builds them
    (src/monte_carlo/monte_carlo.h:199-271)."""
    lib = library()
    d = str(directory).encode()
    nt, npt = C.c_int64(), C.c_int64()
    _check(lib.kmc_read_mesh(d, 0, C.byref(nt), C.byref(npt), *([None] * 8)))
    N = nt.value * npt.value
    out = {k: np.zeros(N) for k in _SITE_KEYS}
    left = np.zeros(N, np.int32)
    right = np.zeros(N, np.int32)
    _check(lib.kmc_read_mesh(d, N, C.byref(nt), C.byref(npt), *[_ptr(out[k], _D) for k in _SITE_KEYS],
                             _ptr(left, _I32), _ptr(right, _I32)))
    out.update(left=left, right=right, chirality=np.zeros(N, np.uint8),
               tube=np.repeat(np.arange(nt.value, dtype=np.int32), npt.value), n=N)
    return out


def trim_sites(net: dict, xlim, ylim, zlim) -> dict:
    """Sites of `net` inside the box, in trim_scats order with repaired links
    (src/monte_carlo/monte_carlo.h:722-782).  Returns a new dict; chirality, flags and tube
    (when present) follow the permutation, source_index holds each kept site's original index."""
    lib = library()
    N = int(np.asarray(net["x"]).size)
    a = {k: np.array(net[k], np.float64) for k in _SITE_KEYS}
    left = np.array(net["left"], np.int32)
    right = np.array(net["right"], np.int32)
    origin = np.zeros(N, np.int64)
    n = C.c_int64(N)
    lim = [np.asarray(v, np.float64) for v in (xlim, ylim, zlim)]
    _check(lib.kmc_trim_sites(C.byref(n), *[_ptr(a[k], _D) for k in _SITE_KEYS], _ptr(left, _I32),
                              _ptr(right, _I32), None, None, _ptr(origin, _I64),
                              *[_ptr(v, _D) for v in lim]))
    M = n.value
    out = {k: a[k][:M].copy() for k in _SITE_KEYS}
    out.update(left=left[:M].copy(), right=right[:M].copy(), n=M, source_index=origin[:M].copy())
    for k in ("chirality", "flags", "tube"):
        if k in net and np.asarray(net[k]).size == N:
            out[k] = np.asarray(net[k])[origin[:M]].copy()
    return out


def kubo_regions(net: dict, n_sections: int):
    """KMC_SITE_SOURCE / KMC_SITE_DRAIN flags of kubo_init's injection region and removal
    domain (src/monte_carlo/monte_carlo.cpp:203-251) and the simulation domain (lo, hi)."""
    a = [_c(net[k], np.float64) for k in ("x", "y", "z")]
    flags = np.zeros(a[0].size, np.uint8)
    lo, hi = np.zeros(3), np.zeros(3)
    _check(library().kmc_kubo_regions(a[0].size, *[_ptr(v, _D) for v in a], int(n_sections),
                                      _ptr(flags, _U8), _ptr(lo, _D), _ptr(hi, _D)))
    return flags, lo, hi


def linspace(start: float, end: float, n: int) -> np.ndarray:
    out = np.zeros(n)
    _check(library().kmc_linspace(start, end, n, _ptr(out, _D)))
    return out


def table_rate(theta, z_shift, a1, a2, rates, q_theta, q_z_shift, q_a1, q_a2, ci=None,
               cj=None):
    """scattering_struct::get_rate (src/monte_carlo/scattering_struct.h:38-52) on the host for
    arrays of geometries: (rates at the nearest nodes, node indices [n][4]).  rates is one table
    (n_theta, n_z, n_a1, n_a2) or n_chir^2 of them stacked in front."""
    g = [_c(v, np.float64) for v in (theta, z_shift, a1, a2)]
    r = _c(rates, np.float64)
    per = g[0].size * g[1].size * g[2].size * g[3].size
    n_chir = int(round(math.sqrt(r.size // per)))
    q = [_c(np.atleast_1d(v), np.float64) for v in (q_theta, q_z_shift, q_a1, q_a2)]
    n = q[0].size
    c1 = None if ci is None else _c(np.broadcast_to(ci, (n,)), np.uint8)
    c2 = None if cj is None else _c(np.broadcast_to(cj, (n,)), np.uint8)
    out = np.zeros(n)
    nodes = np.zeros((n, 4), np.int32)
    _check(library().kmc_table_rate(n_chir, g[0].size, _ptr(g[0], _D), g[1].size, _ptr(g[1], _D),
                                    g[2].size, _ptr(g[2], _D), g[3].size, _ptr(g[3], _D),
                                    _ptr(r, _D), n, _ptr(c1, _U8), _ptr(c2, _U8),
                                    *[_ptr(v, _D) for v in q], _ptr(out, _D), _ptr(nodes, _I32)))
    return out, nodes


def ff_time(gamma, r) -> np.ndarray:
    """ff_time (src/monte_carlo/scatterer.h:74-80), -(1/gamma) log(r / RAND_MAX), with the hop
    kernels' logarithm, evaluated on the host (arrays broadcast)."""
    g, rr = np.broadcast_arrays(np.asarray(gamma, np.float64), np.asarray(r, np.uint32))
    g, rr = _c(g, np.float64), _c(rr, np.uint32)
    out = np.zeros(g.size)
    _check(library().kmc_ff_time(g.size, _ptr(g, _D), _ptr(rr, _U32), _ptr(out, _D)))
    return out.reshape(g.shape)


def forster_table(gamma_0: float, theta, z_shift, a1, a2) -> np.ndarray:
    g = [_c(v, np.float64) for v in (theta, z_shift, a1, a2)]
    out = np.zeros(tuple(v.size for v in g))
    _check(library().kmc_forster_table(gamma_0, g[0].size, _ptr(g[0], _D), g[1].size,
                                       _ptr(g[1], _D), g[2].size, _ptr(g[2], _D), g[3].size,
                                       _ptr(g[3], _D), _ptr(out, _D)))
    return out


def save_rate_table(directory, theta, z_shift, a1, a2, rates, name: str = "scat_table"):
    """Writes one rate table as scattering_struct::save lays it out
    (src/monte_carlo/scattering_struct.h:55-94): <name>.{theta,z_shift,axis_shift_1,
    axis_shift_2,rates}.dat in `directory`, with 17 significant digits so that load_rate_table
    gives back the same doubles.  rates has shape (n_theta, n_z, n_a1, n_a2)."""
    g = [_c(v, np.float64) for v in (theta, z_shift, a1, a2)]
    r = _c(rates, np.float64)
    if r.shape != tuple(v.size for v in g):
        raise ValueError(f"rates shape {r.shape} does not match the axes {tuple(v.size for v in g)}")
    _check(library().kmc_save_rate_table(str(directory).encode(), name.encode(),
                                         *[v.size for v in g], *[_ptr(v, _D) for v in g],
                                         _ptr(r, _D)))


def load_rate_table(directory, name: str = "scat_table"):
    """(theta, z_shift, a1, a2, rates) read from the files save_rate_table (or the reference's
    scattering_struct::save) wrote; rates has shape (n_theta, n_z, n_a1, n_a2)."""
    lib = library()
    d, nm = str(directory).encode(), name.encode()
    dims = np.zeros(4, np.int64)
    _check(lib.kmc_load_rate_table(d, nm, _ptr(dims, _I64), None, None, None, None, None))
    axes = [np.zeros(int(n)) for n in dims]
    rates = np.zeros(tuple(int(n) for n in dims))
    _check(lib.kmc_load_rate_table(d, nm, _ptr(dims, _I64), *[_ptr(v, _D) for v in axes],
                                   _ptr(rates, _D)))
    return (*axes, rates)


def fick_diffusion(population, crossings, area, ymin: float, dy: float, dt: float,
                   steady_from: int = 0, interior_fit: bool = False):
    """Fick's law D = J / |dn/dy| from a contact run: densities n = population / (area dy) and
    currents J = crossings / (interface area dt) (save_population_profile / save_currents,
    monte_carlo.h:525-626) averaged over the steps from steady_from on; J is the mean of |J| at
    the first and the last interface.  The gradient is the reference profiler's
    (visualization/profiler.py:55-65: the density difference of the two end sections over the
    distance of their centres) or, with interior_fit, the least-squares slope over the sections
    between the contacts (the contacts are reset every step, which bends the profile next to
    them).  Returns (D, J, gradient, mean density profile)."""
    pop = np.asarray(population, np.float64)[steady_from:]
    cur = np.asarray(crossings, np.float64)[steady_from:]
    area = np.asarray(area, np.float64)
    S = area.size
    dens = pop / (area[None, :] * dy)
    a_if = (area[:-1] + area[1:]) / 2
    J = cur / (a_if[None, :] * dt)
    avg_j = J.mean(0)
    j = 0.5 * (abs(avg_j[0]) + abs(avg_j[-1]))
    n_avg = dens.mean(0)
    pos = ymin + (np.arange(S) + 0.5) * dy
    if interior_fit:
        if S < 4:
            raise ValueError("interior_fit needs at least two sections between the contacts")
        grad = abs(np.polyfit(pos[1:-1], n_avg[1:-1], 1)[0])
    else:
        grad = abs((n_avg[0] - n_avg[-1]) / (pos[-1] - pos[0]))
    return j / grad, j, grad, n_avg


def estimate_diffusion(reduced: np.ndarray, t_sample_max: float, k_lo: int = 0,
                       k_hi: int | None = None):
    red = _c(reduced, np.float64)
    T, B = red.shape[0], red.shape[1]
    D = np.zeros(3)
    ci = np.zeros(3)
    _check(library().kmc_estimate_diffusion(_ptr(red, _D), T, B, float(t_sample_max), int(k_lo),
                                            int(T if k_hi is None else k_hi), _ptr(D, _D),
                                            _ptr(ci, _D)))
    return D, ci


def default_grids():
    """The table grid of input.json:161-164 (z 1.5..10 nm x11, shifts -10..10 nm x11, theta
    0..180 deg x21), built like create_*_scatt_table (monte_carlo.cpp:65-75)."""
    pi_ref = 3.141592  # src/helper/constants.h:10
    z = linspace(1.5e-9, 10e-9, 11)
    a1 = linspace(-10e-9, 10e-9, 11)
    a2 = linspace(-10e-9, 10e-9, 11)
    th = linspace(0.0, 180.0, 21) * (pi_ref / 180)
    return th, z, a1, a2
