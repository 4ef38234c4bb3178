"""ctypes front end of the CPU oracle (liboracle_kmc.so, built from oracle/kmc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py, as the checker or the CPU baseline.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_kmc.so"

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)
_U32 = C.POINTER(C.c_uint32)
_U64 = C.POINTER(C.c_uint64)


class RunArgs(C.Structure):
    _fields_ = [
        ("n_sites", C.c_int64), ("row_ptr", _I64), ("col", _I32), ("cum", _D), ("pos", _F),
        ("flags", _U8), ("n_src", C.c_int64), ("src", _I32), ("seed", C.c_uint64),
        ("E", C.c_int64), ("site", _I32), ("t_next", _D), ("hops", _U64), ("dsp", _D),
        ("drains", _U32), ("hash", _U64), ("k_rec", _U32),
        ("hop_budget", C.c_int64), ("t_max", C.c_double),
        ("n_bins", C.c_int32), ("t_sample_max", C.c_double), ("rec_disp", _D),
        ("rec_drain", _U32), ("site_count", _U64),
        ("n_trace", C.c_int64), ("tr_max", C.c_int64), ("tr_site", _I32), ("tr_time", _D),
        ("tr_n", _I64), ("net", _P), ("n_threads", C.c_int), ("per_len", C.c_double * 3),
        ("velocity", C.c_double), ("left", _I32), ("right", _I32), ("heading", _U8),
        ("s_off", _D), ("t_last", _D), ("fly_err", C.POINTER(C.c_int32)),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(str(LIB_PATH))
        sig = {
            "orc_philox4x32_10": (None, [_U32, _U32, _U32]),
            "orc_linspace": (C.c_int, [C.c_double, C.c_double, C.c_int, _D]),
            "orc_forster_table": (C.c_int, [C.c_double, C.c_int, _D, C.c_int, _D, C.c_int, _D,
                                            C.c_int, _D, _D]),
            "orc_net_create": (_P, [C.c_int64, _D, _D, _D, _D, _D, _D, _U8, _U8, C.c_double,
                                    C.c_double, C.c_int, _D, _D]),
            "orc_net_free": (None, [_P]),
            "orc_net_cells": (C.c_int, [_P, _I32, _I64, _I64, _I32]),
            "orc_net_set_table": (C.c_int, [_P, C.c_int, C.c_int, _D, C.c_int, _D, C.c_int, _D,
                                            C.c_int, _D, _D]),
            "orc_net_count_pairs": (C.c_int64, [_P, _I64]),
            "orc_net_fill": (C.c_int, [_P, C.c_int, _I64, _D, _I32, _D, _F, _D, _D, _D, _D]),
            "orc_select_index": (C.c_int, [_D, C.c_int, C.c_double]),
            "orc_dwell_time": (C.c_double, [C.c_double, C.c_uint32]),
            "orc_div_rand_max_mismatches": (C.c_int64, [C.c_uint32, C.c_uint32, C.c_int64,
                                                        C.c_uint64]),
            "orc_inject": (C.c_int, [C.c_int64, _I64, _D, C.c_int64, _I32, C.c_uint64, C.c_int64,
                                     _I32, _D, _U64, _D, _U32, _U64, _U32]),
            "orc_run": (C.c_int, [C.POINTER(RunArgs)]),
            "orc_inject_flight": (C.c_int, [C.c_uint64, C.c_int64, _U8, _D, _D]),
            "orc_reduce_samples": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, _U32, _D, _U32,
                                             _D]),
            "orc_max_threads": (C.c_int, []),
            "orc_table_rate": (C.c_int, [C.c_int, C.c_int, _D, C.c_int, _D, C.c_int, _D, C.c_int,
                                         _D, _D, C.c_int64, _U8, _U8, _D, _D, _D, _D, _D, _I32]),
            "orc_contact": (C.c_int, [C.c_int64, _D, _D, _D, _I64, _I32, _D, C.c_uint64, C.c_int,
                                      C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_int64,
                                      C.c_double, _I32, _I32, _F, _D,
                                      _I32, _I32, _I64, _D, _I64, _U32, _I32, _D, _U64, _D,
                                      _U64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(ct)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def philox(ctr, key) -> np.ndarray:
    c = _c(ctr, np.uint32)
    k = _c(key, np.uint32)
    out = np.zeros(4, np.uint32)
    library().orc_philox4x32_10(_p(c, _U32), _p(k, _U32), _p(out, _U32))
    return out


def linspace(a, b, n) -> np.ndarray:
    out = np.zeros(n)
    library().orc_linspace(a, b, n, _p(out, _D))
    return out


def forster_table(gamma_0, th, z, a1, a2) -> np.ndarray:
    g = [_c(v, np.float64) for v in (th, z, a1, a2)]
    out = np.zeros(tuple(v.size for v in g))
    library().orc_forster_table(gamma_0, g[0].size, _p(g[0], _D), g[1].size, _p(g[1], _D),
                                g[2].size, _p(g[2], _D), g[3].size, _p(g[3], _D), _p(out, _D))
    return out


def select_index(cum, dice) -> int:
    c = _c(cum, np.float64)
    return library().orc_select_index(_p(c, _D), c.size, float(dice))


def table_rate(theta, z_shift, a1, a2, rates, q_theta, q_z_shift, q_a1, q_a2, ci=None, cj=None):
    """get_rate of scattering_struct (oracle restatement) for arrays of geometries:
    (rates, node indices [n][4])."""
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
    library().orc_table_rate(n_chir, g[0].size, _p(g[0], _D), g[1].size, _p(g[1], _D), g[2].size,
                             _p(g[2], _D), g[3].size, _p(g[3], _D), _p(r, _D), n, _p(c1, _U8),
                             _p(c2, _U8), *[_p(v, _D) for v in q], _p(out, _D), _p(nodes, _I32))
    return out, nodes


def dwell_time(max_rate, r) -> float:
    return library().orc_dwell_time(float(max_rate), int(r))


class Net:
    """Oracle network: sites, cell list, rate table, CSR builder."""

    def __init__(self, net: dict, cutoff: float, min_distance: float = 0.4e-9,
                 periodic_mask: int = 0, period_lo=(0.0, 0.0, 0.0), period_len=(0.0, 0.0, 0.0)):
        self.lib = library()
        self.arr = {k: _c(net[k], np.float64) for k in ("x", "y", "z", "ox", "oy", "oz")}
        n = self.arr["x"].size
        self.n = n
        self.chir = _c(net.get("chirality", np.zeros(n)), np.uint8)
        self.flags = _c(net.get("flags", np.zeros(n)), np.uint8)
        plo = _c(period_lo, np.float64)
        pln = _c(period_len, np.float64)
        self.period = np.array([pln[a] if periodic_mask & (1 << a) else 0.0 for a in range(3)])
        self.h = self.lib.orc_net_create(
            n, *[_p(self.arr[k], _D) for k in ("x", "y", "z", "ox", "oy", "oz")],
            _p(self.chir, _U8), _p(self.flags, _U8), cutoff, min_distance, periodic_mask,
            _p(plo, _D), _p(pln, _D))
        if not self.h:
            raise RuntimeError("orc_net_create failed")
        self.src = np.nonzero(self.flags & 1)[0].astype(np.int32)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_net_free(self.h)
            self.h = None

    def cells(self):
        dims = np.zeros(3, np.int32)
        nc = C.c_int64()
        self.lib.orc_net_cells(self.h, _p(dims, _I32), C.byref(nc), None, None)
        start = np.zeros(nc.value + 1, np.int64)
        sites = np.zeros(self.n, np.int32)
        self.lib.orc_net_cells(self.h, None, None, _p(start, _I64), _p(sites, _I32))
        return dims, start, sites

    def set_table(self, th, z, a1, a2, rates):
        g = [_c(v, np.float64) for v in (th, z, a1, a2)]
        r = _c(rates, np.float64)
        per = g[0].size * g[1].size * g[2].size * g[3].size
        n_chir = int(round(math.sqrt(r.size // per)))
        self.lib.orc_net_set_table(self.h, n_chir, g[0].size, _p(g[0], _D), g[1].size,
                                   _p(g[1], _D), g[2].size, _p(g[2], _D), g[3].size,
                                   _p(g[3], _D), _p(r, _D))

    def csr(self, mode: int = 0, pair_rates=None, geometry: bool = False) -> dict:
        row = np.zeros(self.n + 1, np.int64)
        nnz = self.lib.orc_net_count_pairs(self.h, _p(row, _I64))
        col = np.zeros(nnz, np.int32)
        cum = np.zeros(nnz)
        disp = np.zeros((nnz, 3), np.float32)
        geo = [np.zeros(nnz) for _ in range(4)] if geometry else [None] * 4
        pr = None if pair_rates is None else _c(pair_rates, np.float64)
        self.lib.orc_net_fill(self.h, mode, _p(row, _I64), _p(pr, _D), _p(col, _I32),
                              _p(cum, _D) if (mode == 0 or pr is not None) else None,
                              _p(disp, _F), *[_p(g, _D) for g in geo])
        out = dict(row_ptr=row, col=col, cum=cum, disp=disp, period=self.period.copy(),
                   site_pos=np.stack([self.arr[k] for k in ("x", "y", "z")], 1).astype(np.float32))
        if geometry:
            out.update(theta=geo[0], z=geo[1], a1=geo[2], a2=geo[3])
        return out


class Excitons:
    """Oracle exciton ensemble over a CSR (or recomputing neighbours per hop)."""

    def __init__(self, csr: dict, flags: np.ndarray, seed: int, n_excitons: int,
                 velocity: float = 0.0, links=None):
        """csr: row_ptr, col, cum, site_pos (FP32 [n][3]) and period (per axis, 0 = open).
        velocity > 0 adds the drift along the tubes; links = (left, right) site links."""
        self.lib = library()
        self.row = _c(csr["row_ptr"], np.int64)
        self.col = _c(csr["col"], np.int32)
        self.cum = _c(csr["cum"], np.float64)
        self.pos = _c(csr["site_pos"], np.float32)
        self.per_len = np.asarray(csr.get("period", (0.0, 0.0, 0.0)), np.float64)
        self.flags = _c(flags, np.uint8)
        self.src = np.nonzero(self.flags & 1)[0].astype(np.int32)
        self.seed = seed
        E = n_excitons
        self.E = E
        self.site = np.zeros(E, np.int32)
        self.t_next = np.zeros(E)
        self.hops = np.zeros(E, np.uint64)
        self.dsp = np.zeros((E, 3))
        self.drains = np.zeros(E, np.uint32)
        self.hash = np.zeros(E, np.uint64)
        self.k_rec = np.zeros(E, np.uint32)
        self.site_count = np.zeros(self.row.size - 1, np.uint64)
        self.n_bins = 0
        self.t_sample_max = 0.0
        self.rec_disp = np.zeros((0, E, 3))
        self.rec_drain = np.zeros((0, E), np.uint32)
        self.n_trace = 0
        self.tr_max = 0
        self.tr_site = np.zeros((0, 0), np.int32)
        self.tr_time = np.zeros((0, 0))
        self.tr_n = np.zeros(0, np.int64)
        rc = self.lib.orc_inject(self.row.size - 1, _p(self.row, _I64), _p(self.cum, _D),
                                 self.src.size, _p(self.src, _I32), seed, E, _p(self.site, _I32),
                                 _p(self.t_next, _D), _p(self.hops, _U64), _p(self.dsp, _D),
                                 _p(self.drains, _U32), _p(self.hash, _U64), _p(self.k_rec, _U32))
        if rc != 0:
            raise RuntimeError("orc_inject: no source sites")
        self.velocity = float(velocity)
        self.heading = np.zeros(E, np.uint8)
        self.s_off = np.zeros(E)
        self.t_last = np.zeros(E)
        self.fly_err = C.c_int32(0)
        if self.velocity > 0.0:
            if links is None:
                raise ValueError("velocity > 0 needs the (left, right) site links")
            self.left = _c(links[0], np.int32)
            self.right = _c(links[1], np.int32)
            self.lib.orc_inject_flight(seed, E, _p(self.heading, _U8), _p(self.s_off, _D),
                                       _p(self.t_last, _D))

    def set_samples(self, n_bins: int, t_sample_max: float):
        self.n_bins, self.t_sample_max = n_bins, t_sample_max
        self.rec_disp = np.zeros((n_bins, self.E, 3))
        self.rec_drain = np.zeros((n_bins, self.E), np.uint32)
        self.k_rec[:] = 0

    def set_trace(self, n_trace: int, max_hops: int):
        self.n_trace, self.tr_max = n_trace, max_hops
        self.tr_site = np.zeros((n_trace, max_hops), np.int32)
        self.tr_time = np.zeros((n_trace, max_hops))
        self.tr_n = np.zeros(n_trace, np.int64)

    def run(self, hops_per_exciton: int, t_max: float = math.inf, count_sites: bool = True,
            net: Net | None = None, n_threads: int = 0):
        a = RunArgs()
        a.n_sites = self.row.size - 1
        a.row_ptr, a.col, a.cum = _p(self.row, _I64), _p(self.col, _I32), _p(self.cum, _D)
        a.pos, a.flags = _p(self.pos, _F), _p(self.flags, _U8)
        for k in range(3):
            a.per_len[k] = float(self.per_len[k])
        a.n_src, a.src, a.seed = self.src.size, _p(self.src, _I32), self.seed
        a.E = self.E
        a.site, a.t_next, a.hops = _p(self.site, _I32), _p(self.t_next, _D), _p(self.hops, _U64)
        a.dsp, a.drains, a.hash = _p(self.dsp, _D), _p(self.drains, _U32), _p(self.hash, _U64)
        a.k_rec = _p(self.k_rec, _U32)
        a.hop_budget, a.t_max = int(hops_per_exciton), float(t_max)
        a.n_bins, a.t_sample_max = self.n_bins, self.t_sample_max
        a.rec_disp, a.rec_drain = _p(self.rec_disp, _D), _p(self.rec_drain, _U32)
        a.site_count = _p(self.site_count, _U64) if count_sites else None
        a.n_trace, a.tr_max = self.n_trace, self.tr_max
        a.tr_site, a.tr_time, a.tr_n = (_p(self.tr_site, _I32), _p(self.tr_time, _D),
                                        _p(self.tr_n, _I64))
        if self.n_trace == 0:
            a.tr_site = None
            a.tr_n = None
        a.net = net.h if net is not None else None
        a.n_threads = n_threads
        a.velocity = self.velocity
        if self.velocity > 0.0:
            a.left, a.right = _p(self.left, _I32), _p(self.right, _I32)
            a.heading, a.s_off, a.t_last = (_p(self.heading, _U8), _p(self.s_off, _D),
                                            _p(self.t_last, _D))
            a.fly_err = C.pointer(self.fly_err)
        self.lib.orc_run(C.byref(a))
        if self.fly_err.value:
            raise RuntimeError("oracle: a flight passed more than 2^22 sites")

    def reduce_samples(self, n_batches: int) -> np.ndarray:
        out = np.zeros((self.n_bins, n_batches, 5))
        self.lib.orc_reduce_samples(self.E, self.n_bins, n_batches, _p(self.k_rec, _U32),
                                    _p(self.rec_disp, _D), _p(self.rec_drain, _U32), _p(out, _D))
        return out


# ---------------------------------------------------------------------------------------------
# exact transfer rate (literal first_order)
# ---------------------------------------------------------------------------------------------
class _Tube(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("Nu", "L", "nk", "n_mu", "nr", "nk_cm", "nk_c", "ik_cm0")] + \
               [("energy", _D), ("psi", _D), ("idx", _I32), ("wf", _D),
                ("dk_l", C.c_double * 2), ("aCC", C.c_double * 2), ("t_vec", C.c_double * 2),
                ("t_vec_3d", C.c_double * 3), ("Au", C.c_double), ("radius", C.c_double),
                ("length_m", C.c_double), ("pos_u_2d", _D), ("pos_u_3d", _D)]


def tube_from_model(model) -> dict:
    """Arrays of a product CntModel (cnt_film_monte_carlo_amd.exciton) for the oracle."""
    i = model.info
    psi, idx = model.exciton_states()
    _, wf = model.electrons()
    pa, _, p3 = model.atoms()
    return dict(Nu=i.Nu, L=i.length_cells, nk=i.nk, n_mu=i.n_mu, nr=i.nk_relev, nk_cm=i.nk_cm,
                nk_c=i.nk_c, ik_cm0=i.ik_cm0, energy=np.ascontiguousarray(model.dispersion()),
                psi=np.ascontiguousarray(psi), idx=np.ascontiguousarray(idx, np.int32),
                wf=np.ascontiguousarray(wf), dk_l=tuple(i.dk_l), aCC=tuple(i.aCC),
                t_vec=tuple(i.t_vec), t_vec_3d=(0.0, i.t_vec[1], 0.0), Au=i.Au, radius=i.radius,
                length_m=i.length_m, pos_u_2d=np.ascontiguousarray(pa),
                pos_u_3d=np.ascontiguousarray(p3))


def _tube_struct(t: dict) -> _Tube:
    s = _Tube()
    for k in ("Nu", "L", "nk", "n_mu", "nr", "nk_cm", "nk_c", "ik_cm0"):
        setattr(s, k, int(t[k]))
    s.energy = t["energy"].ctypes.data_as(_D)
    s.psi = t["psi"].ctypes.data_as(_D)
    s.idx = t["idx"].ctypes.data_as(_I32)
    s.wf = t["wf"].ctypes.data_as(_D)
    for k in ("dk_l", "aCC", "t_vec", "t_vec_3d"):
        arr = getattr(s, k)
        for q, v in enumerate(t[k]):
            arr[q] = v
    s.Au, s.radius, s.length_m = t["Au"], t["radius"], t["length_m"]
    s.pos_u_2d = t["pos_u_2d"].ctypes.data_as(_D)
    s.pos_u_3d = t["pos_u_3d"].ctypes.data_as(_D)
    return s


def first_order(donor: dict, acceptor: dict, temperature: float, broadening: float, z_shift: float,
                axis_shift_1: float, axis_shift_2: float, theta: float):
    """Literal first_order on the CPU; returns (rate, number of matched pairs)."""
    lib = library()
    fn = lib.orc_first_order
    fn.restype = C.c_double
    fn.argtypes = [C.POINTER(_Tube), C.POINTER(_Tube), C.c_double, C.c_double, C.c_double,
                   C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int)]
    sd, sa = _tube_struct(donor), _tube_struct(acceptor)
    npairs = C.c_int(0)
    r = fn(C.byref(sd), C.byref(sa), temperature, broadening, z_shift, axis_shift_1, axis_shift_2,
           theta, C.byref(npairs))
    return r, npairs.value


def contact(net: dict, csr: dict, seed: int, n_seg: int, c1_pop: int, c2_pop: int,
            capacity: int, dt: float, n_steps: int, velocity: float = 0.0) -> dict:
    """Contact (steady-state) mode over the oracle CSR (orc_contact: init, step, save_metrics,
    repopulate_contacts of src/monte_carlo/monte_carlo.h, written literally): per step the
    population [n_steps, n_seg], the interface crossings [n_steps, n_seg - 1] and the live count;
    the section areas; the final particles (id, site, t_next, hops, y) in list order.  velocity > 0
    adds the drift along the tubes (particle::fly) over net's left / right links, the FP32
    csr["site_pos"] and the periods csr["period"]."""
    lib = library()
    velocity = float(velocity)
    if velocity > 0.0:
        left, right = _c(net["left"], np.int32), _c(net["right"], np.int32)
        spos = _c(csr["site_pos"], np.float32)
        per = _c(np.asarray(csr.get("period", (0.0, 0.0, 0.0)), np.float64), np.float64)
        fly = (velocity, _p(left, _I32), _p(right, _I32), _p(spos, _F), _p(per, _D))
    else:
        fly = (0.0, None, None, None, None)
    pos = [_c(net[k], np.float64) for k in ("x", "y", "z")]
    row = _c(csr["row_ptr"], np.int64)
    col = _c(csr["col"], np.int32)
    cum = _c(csr["cum"], np.float64)
    pop = np.zeros((n_steps, n_seg), np.int32)
    cur = np.zeros((n_steps, n_seg - 1), np.int32)
    live = np.zeros(n_steps, np.int64)
    area = np.zeros(n_seg)
    n_final = np.zeros(1, np.int64)
    pid = np.zeros(capacity, np.uint32)
    site = np.zeros(capacity, np.int32)
    t = np.zeros(capacity)
    hops = np.zeros(capacity, np.uint64)
    yy = np.zeros(capacity)
    made = np.zeros(1, np.uint64)
    rc = lib.orc_contact(pos[0].size, *[_p(v, _D) for v in pos], _p(row, _I64), _p(col, _I32),
                         _p(cum, _D), seed, n_seg, c1_pop, c2_pop, capacity, dt, n_steps, *fly,
                         _p(pop, _I32), _p(cur, _I32), _p(live, _I64), _p(area, _D),
                         _p(n_final, _I64), _p(pid, _U32), _p(site, _I32), _p(t, _D),
                         _p(hops, _U64), _p(yy, _D), _p(made, _U64))
    if rc != 0:
        raise RuntimeError(f"orc_contact failed ({rc})")
    n = int(n_final[0])
    return dict(population=pop, crossings=cur, n_live=live, area=area, id=pid[:n],
                site=site[:n], t_next=t[:n], hops=hops[:n], y=yy[:n], hops_made=int(made[0]))

