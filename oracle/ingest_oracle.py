"""CPU restatement of the reference's site ingest, for tests only (test infrastructure: imported by
tests/, never by the product).

* trim: trim_scats (src/monte_carlo/monte_carlo.h:722-782), the swap-with-last removal with its
  link repairs, as a literal loop over Python lists;
* kubo_regions: find_simulation_domain (monte_carlo.h:328-340), injection_region and
  get_removal_domain (src/monte_carlo/monte_carlo.cpp:203-251);
* mesh_sites: create_scatterers (monte_carlo.h:199-271) applied to matrices already in memory
  (the file format is pinned by a write/read round trip in the tests).
"""
from __future__ import annotations

import numpy as np

KEYS = ("x", "y", "z", "ox", "oy", "oz")


def mesh_sites(pos_nm, orient):
    """pos_nm, orient: three [tubes][points] matrices each -> site dict (positions in m)."""
    rows, cols = np.asarray(pos_nm[0]).shape
    out = {k: [] for k in KEYS}
    left, right = [], []
    for i in range(rows):
        for j in range(cols):
            n = i * cols + j
            for k, name in enumerate(("x", "y", "z")):
                out[name].append(float(pos_nm[k][i][j]) * 1.e-9)
            for k, name in enumerate(("ox", "oy", "oz")):
                out[name].append(float(orient[k][i][j]))
            left.append(n - 1 if j > 0 else -1)
            right.append(n + 1 if j + 1 < cols else -1)
    res = {k: np.array(v) for k, v in out.items()}
    res["left"] = np.array(left, np.int32)
    res["right"] = np.array(right, np.int32)
    return res


def trim(net: dict, xlim, ylim, zlim) -> dict:
    """trim_scats on a site dict; returns the kept sites in the reference's slot order, their
    links, and origin (the original index of every slot)."""
    n = len(net["x"])
    sites = [dict(x=float(net["x"][i]), y=float(net["y"][i]), z=float(net["z"][i]),
                  ox=float(net["ox"][i]), oy=float(net["oy"][i]), oz=float(net["oz"][i]),
                  left=int(net["left"][i]), right=int(net["right"][i]), origin=i)
             for i in range(n)]

    def swap_scatterers(i, j):
        i_left, i_right = sites[i]["left"], sites[i]["right"]
        j_left, j_right = sites[j]["left"], sites[j]["right"]
        new_i, new_j = j, i
        if i_left > -1:
            sites[i_left]["right"] = new_i
        if i_right > -1:
            sites[i_right]["left"] = new_i
        if j_left > -1:
            sites[j_left]["right"] = new_j
        if j_right > -1:
            sites[j_right]["left"] = new_j
        sites[i], sites[j] = sites[j], sites[i]

    j = n
    i = 0
    while i < j:
        s = sites[i]
        if (s["x"] < xlim[0] or s["y"] < ylim[0] or s["z"] < zlim[0] or
                s["x"] > xlim[1] or s["y"] > ylim[1] or s["z"] > zlim[1]):
            j -= 1
            swap_scatterers(i, j)
            if sites[j]["left"] > -1:
                sites[sites[j]["left"]]["right"] = -1
            if sites[j]["right"] > -1:
                sites[sites[j]["right"]]["left"] = -1
        else:
            i += 1
    kept = sites[:j]
    out = {k: np.array([s[k] for s in kept], float) for k in KEYS}
    out["left"] = np.array([s["left"] for s in kept], np.int32)
    out["right"] = np.array([s["right"] for s in kept], np.int32)
    out["origin"] = np.array([s["origin"] for s in kept], np.int64)
    return out


def kubo_regions(x, y, z, n_sections: int):
    """(flags, lo, hi): bit 1 = injection region, bit 2 = outside the removal domain."""
    assert n_sections > 1 and n_sections % 2 == 1
    p = [list(map(float, x)), list(map(float, y)), list(map(float, z))]
    lo = [p[k][0] for k in range(3)]
    hi = [p[k][0] for k in range(3)]
    for i in range(len(p[0])):
        for k in range(3):
            lo[k] = p[k][i] if lo[k] > p[k][i] else lo[k]
            hi[k] = p[k][i] if hi[k] < p[k][i] else hi[k]
    edges = []
    for k in range(3):
        d = (hi[k] - lo[k]) / float(n_sections)
        edges.append([float(i) * d + lo[k] for i in range(n_sections + 1)])
    c = n_sections // 2
    flags = np.zeros(len(p[0]), np.uint8)
    for i in range(len(p[0])):
        inj = all(edges[k][c] <= p[k][i] <= edges[k][c + 1] for k in range(3))
        out = any(p[k][i] < edges[k][1] or edges[k][n_sections - 1] < p[k][i] for k in range(3))
        flags[i] = (1 if inj else 0) | (2 if out else 0)
    return flags, np.array(lo), np.array(hi)
