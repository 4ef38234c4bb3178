"""Synthetic CNT-film workloads: generator parameters, contacts and rate tables.

Each workload is a seeded network of straight tubes (kmc_generate_network) plus the
contact flags of the film: injection (source) sites at x <= x_source and drain sites at
x >= x_drain, with y periodic.  Rates come from a nearest-node table on the reference grid
(input.json "zshift [m]", "axis shift 1/2 [m]", "theta [degrees]").
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import kmc

NM = 1e-9


@dataclass
class Workload:
    name: str
    net: kmc.NetworkParams
    cutoff: float
    x_source: float            # sites with x <= x_source inject
    x_drain: float             # sites with x >= x_drain drain
    period_y: tuple            # (lo, len) of the periodic y axis
    n_excitons: int
    hops_per_exciton: int
    temperature: float = 300.0
    chiralities: tuple = ((7, 5),)
    extra: dict = field(default_factory=dict)


def film_small(seed: int = 1001) -> Workload:
    """250 straight 200 nm CNTs, random in-plane angle, centres in 500x500x50 nm, 10 nm
    segments (5,000 sites), chirality (7,5), cutoff 15 nm, 1,000 excitons x 10,000 hops."""
    net = kmc.NetworkParams(n_tubes=250, len_min=200 * NM, len_max=200 * NM,
                            box_lo=(0.0, 0.0, 0.0), box_hi=(500 * NM, 500 * NM, 50 * NM),
                            site_spacing=10 * NM, n_chir=1, wrap_mask=kmc.PERIODIC_Y, seed=seed)
    return Workload("film_5k_sites", net, cutoff=15 * NM, x_source=0.0, x_drain=500 * NM,
                    period_y=(0.0, 500 * NM), n_excitons=1000, hops_per_exciton=10_000)


def film_large(seed: int = 1002) -> Workload:
    """5,000 CNTs of length U[0.5, 2] um in a 5x5x0.2 um box, 10 nm segments (~625k sites),
    chiralities (7,5), (8,6), (9,7) at 1/3 each, cutoff 20 nm, 1M excitons x 1e5 hops."""
    net = kmc.NetworkParams(n_tubes=5000, len_min=500 * NM, len_max=2000 * NM,
                            box_lo=(0.0, 0.0, 0.0), box_hi=(5000 * NM, 5000 * NM, 200 * NM),
                            site_spacing=10 * NM, n_chir=3, wrap_mask=kmc.PERIODIC_Y, seed=seed)
    return Workload("film_625k_sites", net, cutoff=20 * NM, x_source=0.0, x_drain=5000 * NM,
                    period_y=(0.0, 5000 * NM), n_excitons=1_000_000, hops_per_exciton=100_000,
                    chiralities=((7, 5), (8, 6), (9, 7)))


def film_xlarge(seed: int = 1005) -> Workload:
    """80,000 CNTs of 200 nm in a 20x20x0.5 um box, 10 nm segments (1.6M sites), chiralities
    (7,5), (8,6), (9,7) at 1/3 each, cutoff 20 nm, y periodic, 16M excitons x 1e5 hops."""
    net = kmc.NetworkParams(n_tubes=80_000, len_min=200 * NM, len_max=200 * NM,
                            box_lo=(0.0, 0.0, 0.0), box_hi=(20_000 * NM, 20_000 * NM, 500 * NM),
                            site_spacing=10 * NM, n_chir=3, wrap_mask=kmc.PERIODIC_Y, seed=seed)
    return Workload("film_1600k_sites", net, cutoff=20 * NM, x_source=0.0, x_drain=20_000 * NM,
                    period_y=(0.0, 20_000 * NM), n_excitons=16_000_000, hops_per_exciton=100_000,
                    chiralities=((7, 5), (8, 6), (9, 7)))


def contact_flags(x: np.ndarray, x_source: float, x_drain: float) -> np.ndarray:
    """KMC_SITE_SOURCE for x <= x_source, KMC_SITE_DRAIN for x >= x_drain."""
    f = np.zeros(x.size, np.uint8)
    f[x <= x_source] |= kmc.SITE_SOURCE
    f[x >= x_drain] |= kmc.SITE_DRAIN
    return f


def forster_tables(n_chir: int, gamma_0: float = 1e13):
    """Wong/Förster closed-form table (rate type "wong", monte_carlo.cpp:55-57) on the default
    grid, replicated for every chirality pair."""
    th, z, a1, a2 = kmc.default_grids()
    one = kmc.forster_table(gamma_0, th, z, a1, a2)
    rates = np.broadcast_to(one, (n_chir * n_chir,) + one.shape).copy()
    return th, z, a1, a2, rates


def build_network(w: Workload) -> dict:
    """Generate the sites of a workload and its contact flags."""
    net = kmc.generate_network(w.net)
    net["flags"] = contact_flags(net["x"], w.x_source, w.x_drain)
    return net


def make_context(w: Workload, net: dict, seed: int = 7, device: int = 0) -> kmc.Context:
    cfg = kmc.Config(cutoff=w.cutoff, min_distance=0.4e-9, periodic_mask=kmc.PERIODIC_Y,
                     period_lo=(0.0, w.period_y[0], 0.0), period_len=(0.0, w.period_y[1], 0.0),
                     seed=seed, device=device)
    ctx = kmc.Context(cfg)
    ctx.set_sites(net["x"], net["y"], net["z"], net["ox"], net["oy"], net["oz"], net["left"],
                  net["right"], net["chirality"], net["flags"])
    return ctx


def ensemble(n_networks: int = 256, seed: int = 3003, n_min: int = 5_000, n_max: int = 200_000,
             t_range=(150.0, 400.0)) -> list:
    """Independent networks by the rules of film_small with site counts log-uniform in
    [n_min, n_max]: 200 nm (7,5) tubes of 20 sites at the density of film_small (the film's x
    extent grows with the tube count, y stays 500 nm periodic, 50 nm thick), excitons = 0.2 x
    sites, temperatures uniform in t_range.  Returns Workloads with `temperature` set and the
    network seed distinct per member."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_networks):
        n_sites = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        n_tubes = max(1, int(round(n_sites / 20)))
        lx = 500 * NM * n_tubes / 250
        net = kmc.NetworkParams(n_tubes=n_tubes, len_min=200 * NM, len_max=200 * NM,
                                box_lo=(0.0, 0.0, 0.0), box_hi=(lx, 500 * NM, 50 * NM),
                                site_spacing=10 * NM, n_chir=1, wrap_mask=kmc.PERIODIC_Y,
                                seed=seed * 1000 + k)
        out.append(Workload(f"ensemble_{k}", net, cutoff=15 * NM, x_source=0.0, x_drain=lx,
                            period_y=(0.0, 500 * NM), n_excitons=max(1, int(0.2 * 20 * n_tubes)),
                            hops_per_exciton=10_000,
                            temperature=float(rng.uniform(*t_range))))
    return out
