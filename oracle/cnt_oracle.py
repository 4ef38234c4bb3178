"""Test infrastructure, not product code: a numpy restatement of the CNT model's two large sums,
the Ohno-potential transform v(q) (cnt::calculate_vq, src/exciton_transfer/cnt.cpp:520-631) and
the polarization PI(q) (cnt::calculate_polarization, cnt.cpp:634-732), over the q grid of
calculate_exciton_dispersion (cnt.cpp:1070-1074).  Inputs are the model's own atoms, lattice
vectors and K2-extended bands (pinned elsewhere by closed forms); numpy's complex exponential and
pairwise sums make it an independent check of both the host and the device sums."""
from __future__ import annotations

import numpy as np

# src/exciton_transfer/constants.h:10-18, in the reference's expressions
PI_TRUNC = 3.141592
EV = 1.6 * 10.0 ** -19.0
EPS0 = 8.85 * 10.0 ** -12
Q0 = 1.6 * 10.0 ** -19
UPP = 11.3 * EV                      # cnt.h:31 (Ohno parameter)


def _q_grid(info):
    iq = np.arange(-(info.nk - 1), info.nk)
    mu = np.arange(-(info.n_mu - 1), info.n_mu)
    dk = np.array(info.dk_l[:])
    k1 = np.array(info.K1[:])
    # q = iq dk_l + mu K1 (cnt.cpp:588)
    qx = iq[:, None] * dk[0] + mu[None, :] * k1[0]
    qy = iq[:, None] * dk[1] + mu[None, :] * k1[1]
    return qx, qy


def relative_positions(info, pos_a, pos_b):
    """rel_pos [4][Nu ncells][2] of cnt.cpp:533-568: sublattice pairs aa, ab, ba, bb relative to
    atom 0, wrapped to x <= |Ch| / 2, repeated over an odd number of unit cells along T."""
    ncells = info.length_cells + (1 - info.length_cells % 2)
    ch_x = info.ch_vec[0]
    t = np.array(info.t_vec[:])
    rel = []
    for p, o in ((pos_a, pos_a[0]), (pos_a, pos_b[0]), (pos_b, pos_a[0]), (pos_b, pos_b[0])):
        d = p - o
        d[:, 0] = np.where(d[:, 0] > ch_x / 2, d[:, 0] - ch_x, d[:, 0])
        half = ncells // 2
        cells = np.arange(-half, half + 1)
        r = d[None, :, :] + cells[:, None, None] * t[None, None, :]
        rel.append(r.reshape(-1, 2))
    return np.stack(rel), ncells


def vq(info, pos_a, pos_b, chunk=64):
    """v(q) [2 nk - 1][2 n_mu - 1][4] complex."""
    rel, ncells = relative_positions(info, pos_a, pos_b)
    coeff = (4.0 * PI_TRUNC * EPS0 * UPP / Q0 / Q0) ** 2
    den = UPP / np.sqrt(coeff * (rel[..., 0] ** 2 + rel[..., 1] ** 2) + 1)      # [4][natom]
    qx, qy = _q_grid(info)
    out = np.zeros(qx.shape + (4,), np.complex128)
    flat_x, flat_y = qx.ravel(), qy.ravel()
    res = np.zeros((flat_x.size, 4), np.complex128)
    for s in range(4):
        for c in range(0, flat_x.size, chunk):
            ph = flat_x[c:c + chunk, None] * rel[s, None, :, 0] + flat_y[c:c + chunk, None] * rel[s, None, :, 1]
            res[c:c + chunk, s] = (np.exp(1j * ph) * den[s][None, :]).sum(1)
    out[...] = res.reshape(out.shape)
    return out / (2 * info.Nu * ncells)


def polarization(info, energy, wf):
    """PI(q) [2 nk - 1][2 n_mu - 1]; energy [2][nk][n_mu] (0 valence, 1 conduction), wf
    [n_mu][nk][band][atom]."""
    nk, n_mu = info.nk, info.n_mu
    nk_k1 = nk * info.Q // info.Nu
    nkm = nk_k1 * info.M
    ik = np.repeat(np.arange(nk), n_mu)
    mu_k = np.tile(np.arange(n_mu), nk)
    out = np.zeros((2 * nk - 1, 2 * n_mu - 1))
    for a, iq in enumerate(range(-(nk - 1), nk)):
        for b, mu_q in enumerate(range(-(n_mu - 1), n_mu)):
            # get_kq (cnt.cpp:650-667): mu folded into [0, n_mu) moving k by nk_K1 M, k mod nk
            mu_kq = mu_k + mu_q
            wraps = np.floor_divide(mu_kq, n_mu)
            mu_kq = mu_kq - wraps * n_mu
            ikq = np.mod(ik + iq + wraps * nkm, nk)
            v_k, c_k = wf[mu_k, ik, 0, :], wf[mu_k, ik, 1, :]
            v_kq, c_kq = wf[mu_kq, ikq, 0, :], wf[mu_kq, ikq, 1, :]
            d1 = (np.conj(v_k) * c_kq).sum(1)
            d2 = (np.conj(c_k) * v_kq).sum(1)
            t = (np.abs(d1) ** 2 / (energy[1, ikq, mu_kq] - energy[0, ik, mu_k]) +
                 np.abs(d2) ** 2 / (energy[1, ik, mu_k] - energy[0, ikq, mu_kq]))
            out[a, b] = 2 * t.sum()
    return out
