"""ctypes bindings of the CNT exciton model and the exact transfer rate (include/cntkmc.h).

CntModel wraps kmc_cnt (cnt::calculate_exciton_dispersion); Transfer wraps kmc_transfer
(exciton_transfer::first_order on the device).  Energies are in joules, lengths in metres.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import kmc

EV = 1.6e-19          # src/exciton_transfer/constants.h:12
BROADENING = 4.0e-3 * EV  # exciton_transfer.h:42


class _CntInfo(C.Structure):
    _fields_ = ([(k, C.c_int32) for k in ("n", "m", "length_cells", "Nu", "Q", "M", "t1", "t2",
                                          "nk", "n_mu", "nk_relev", "nk_cm", "nk_c", "ik_cm0")]
                + [(k, C.c_double) for k in ("radius", "length_m", "Au", "a_l")]
                + [(k, C.c_double * 2) for k in ("ch_vec", "t_vec", "aCC", "dk_l", "K1", "K2")]
                + [("e_min", C.c_double)])


class _TransferInfo(C.Structure):
    _fields_ = ([(k, C.c_int32) for k in ("n_donor_states", "n_acceptor_states", "n_pairs",
                                          "n_mom_donor", "n_mom_acceptor", "n_terms",
                                          "n_atoms_donor", "n_atoms_acceptor")]
                + [("partition_Z", C.c_double), ("min_energy", C.c_double)])


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_SIGS = {
    "kmc_cnt_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_P)]),
    "kmc_cnt_create_on_device": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                           C.c_int32, C.POINTER(_P)]),
    "kmc_cnt_destroy": (C.c_int, [_P]),
    "kmc_cnt_last_error": (C.c_char_p, []),
    "kmc_cnt_get_info": (C.c_int, [_P, C.POINTER(_CntInfo)]),
    "kmc_cnt_dispersion": (C.c_int, [_P, _D]),
    "kmc_cnt_electrons": (C.c_int, [_P, _D, _D]),
    "kmc_cnt_interaction": (C.c_int, [_P, _D, _D, _D]),
    "kmc_cnt_exciton_states": (C.c_int, [_P, _D, _I32]),
    "kmc_cnt_singlet_kernel": (C.c_int, [_P, C.c_int32, _D]),
    "kmc_cnt_atoms": (C.c_int, [_P, _D, _D, _D]),
    "kmc_hermitian_eig": (C.c_int, [C.c_int32, _D, _D, _D]),
    "kmc_transfer_create": (C.c_int, [C.c_int32, _P, _P, C.c_double, C.c_double, C.POINTER(_P)]),
    "kmc_transfer_destroy": (C.c_int, [_P]),
    "kmc_transfer_last_error": (C.c_char_p, []),
    "kmc_transfer_get_info": (C.c_int, [_P, C.POINTER(_TransferInfo)]),
    "kmc_transfer_terms": (C.c_int, [_P, _I32, _I32, _D]),
    "kmc_transfer_set_clamp": (C.c_int, [_P, _D, _D]),
    "kmc_transfer_rates": (C.c_int, [_P, C.c_int64, _D, _D, _D, _D, _D, _D]),
    "kmc_transfer_rates_device": (C.c_int, [_P, C.c_int64, _P, _P, _P, _P, _P, _P]),
    "kmc_transfer_table": (C.c_int, [_P, C.c_int32, _D, C.c_int32, _D, C.c_int32, _D, C.c_int32, _D,
                                     _D, _D]),
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


def hermitian_eig(a: np.ndarray):
    """Eigen-decomposition of a Hermitian matrix with the library's solver."""
    a = np.asfortranarray(a, dtype=np.complex128)
    n = a.shape[0]
    ev = np.zeros(n)
    vec = np.zeros((n, n), np.complex128, order="F")
    rc = library().kmc_hermitian_eig(n, a.ctypes.data_as(_D), _p(ev), vec.ctypes.data_as(_D))
    if rc < 0:
        raise kmc.KmcError(rc, library().kmc_cnt_last_error().decode())
    return ev, vec


class CntModel:
    """(n, m) tube of `length` unit cells: A2 singlet exciton dispersion and wavefunctions.
    device (a GPU index) takes the Coulomb transform v(q) and the polarization PI(q), the bulk
    of the build, on that GPU (kmc_cnt_create_on_device); None builds on the host only."""

    def __init__(self, n: int, m: int, length: int = 10, threads: int = 0, device=None):
        lib = library()
        h = _P()
        if device is None:
            rc = lib.kmc_cnt_create(n, m, length, threads, C.byref(h))
        else:
            rc = lib.kmc_cnt_create_on_device(n, m, length, threads, int(device), C.byref(h))
        if rc < 0:
            raise kmc.KmcError(rc, lib.kmc_cnt_last_error().decode())
        self.h = h
        self.lib = lib
        info = _CntInfo()
        lib.kmc_cnt_get_info(h, C.byref(info))
        self.info = info

    def close(self):
        if getattr(self, "h", None):
            self.lib.kmc_cnt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def chirality(self):
        return (self.info.n, self.info.m)

    def dispersion(self) -> np.ndarray:
        i = self.info
        e = np.zeros((i.nk_cm, i.nk_relev))
        self.lib.kmc_cnt_dispersion(self.h, _p(e))
        return e

    def electrons(self):
        i = self.info
        e = np.zeros((2, i.nk, i.n_mu))
        wf = np.zeros((i.n_mu, i.nk, 2, 2), np.complex128)
        self.lib.kmc_cnt_electrons(self.h, _p(e), wf.ctypes.data_as(_D))
        return e, wf

    def interaction(self):
        i = self.info
        nq, nm = 2 * i.nk - 1, 2 * i.n_mu - 1
        vq = np.zeros((nq, nm, 4), np.complex128)
        PI = np.zeros((nq, nm))
        eps = np.zeros((nq, nm))
        self.lib.kmc_cnt_interaction(self.h, vq.ctypes.data_as(_D), _p(PI), _p(eps))
        return vq, PI, eps

    def exciton_states(self):
        i = self.info
        psi = np.zeros((i.nk_cm, i.nk_relev, i.nk_c), np.complex128)
        idx = np.zeros((i.nk_cm, i.nk_c, 4), np.int32)
        self.lib.kmc_cnt_exciton_states(self.h, psi.ctypes.data_as(_D), _p(idx, _I32))
        return psi, idx

    def singlet_kernel(self, ik_cm_idx: int) -> np.ndarray:
        n = self.info.nk_relev
        k = np.zeros((n, n), np.complex128, order="F")
        rc = self.lib.kmc_cnt_singlet_kernel(self.h, ik_cm_idx, k.ctypes.data_as(_D))
        if rc < 0:
            raise kmc.KmcError(rc, self.lib.kmc_cnt_last_error().decode())
        return k

    def atoms(self):
        nu = self.info.Nu
        pa = np.zeros((nu, 2))
        pb = np.zeros((nu, 2))
        p3 = np.zeros((nu, 3))
        self.lib.kmc_cnt_atoms(self.h, _p(pa), _p(pb), _p(p3))
        return pa, pb, p3


class Transfer:
    """Exact first-order transfer rate between a donor and an acceptor tube."""

    def __init__(self, donor: CntModel, acceptor: CntModel, temperature: float = 300.0,
                 broadening: float = BROADENING, device: int = 0):
        lib = library()
        h = _P()
        rc = lib.kmc_transfer_create(device, donor.h, acceptor.h, temperature, broadening,
                                     C.byref(h))
        if rc < 0:
            raise kmc.KmcError(rc, lib.kmc_transfer_last_error().decode())
        self.h = h
        self.lib = lib
        self.donor, self.acceptor = donor, acceptor   # keep the models alive
        info = _TransferInfo()
        lib.kmc_transfer_get_info(h, C.byref(info))
        self.info = info

    def close(self):
        if getattr(self, "h", None):
            self.lib.kmc_transfer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def terms(self):
        """(donor ik_cm, acceptor ik_cm, C_ab) of every nonzero term."""
        n = self.info.n_terms
        ka = np.zeros(n, np.int32)
        kb = np.zeros(n, np.int32)
        c = np.zeros(n)
        self.lib.kmc_transfer_terms(self.h, _p(ka, _I32), _p(kb, _I32), _p(c))
        return ka, kb, c

    def set_clamp(self, lo, hi):
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        rc = self.lib.kmc_transfer_set_clamp(self.h, _p(lo), _p(hi))
        if rc < 0:
            raise kmc.KmcError(rc, self.lib.kmc_transfer_last_error().decode())

    def rates(self, theta, z, a1, a2):
        g = [np.ascontiguousarray(v, np.float64).ravel() for v in (theta, z, a1, a2)]
        n = g[0].size
        out = np.zeros(n)
        ms = C.c_double(0.0)
        rc = self.lib.kmc_transfer_rates(self.h, n, *[_p(v) for v in g], _p(out), C.byref(ms))
        if rc < 0:
            raise kmc.KmcError(rc, self.lib.kmc_transfer_last_error().decode())
        self.last_ms = ms.value
        return out

    def table(self, theta, z, a1, a2):
        """Exact 4-D table [n_theta][n_z][n_a1][n_a2] on the grid (create_davoody_scatt_table).
        Sets self.last_ms to the wall time of the build."""
        g = [np.ascontiguousarray(v, np.float64).ravel() for v in (theta, z, a1, a2)]
        out = np.zeros([v.size for v in g])
        ms = C.c_double(0.0)
        rc = self.lib.kmc_transfer_table(self.h, g[0].size, _p(g[0]), g[1].size, _p(g[1]), g[2].size,
                                         _p(g[2]), g[3].size, _p(g[3]), _p(out), C.byref(ms))
        if rc < 0:
            raise kmc.KmcError(rc, self.lib.kmc_transfer_last_error().decode())
        self.last_ms = ms.value
        return out
