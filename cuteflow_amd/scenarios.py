"""Seeded synthetic workloads of the river-reach family (no external data).

- reach(): one river reach with a floodplain at any cell count, the same
  construction as the 2,048,000-cell reach (20 km x 2 km, 3200 x 320 quads
  split into triangles, 6.25 m cells), either at the same cell size (the
  reach grows with the cell count) or on the same 20 km x 2 km reach with
  larger cells;
- ensemble(): the batched ensemble: the 20 km x 2 km reach meshed at uneven
  sizes drawn uniformly in [nmin, nmax] cells, each member with its own peak
  inflow and Manning n in 8 zones (4 reaches along the flow x
  channel/floodplain);
- calibration(): the zonal calibration reach: 8 zones, gauge cells in the
  channel, a hidden true n per zone and the start 20 % above it;
- calibration_batch(): many calibration scenarios (flood events) on one
  reach, each with its own peak inflow, hidden zonal n and gauges;
- field_ensemble(): members with a per-cell log-normal Manning field and
  their own inflow hydrographs (the floodplain wets and dries as the flood
  passes).
Choices the configurations leave open are fixed here (zone layout, gauge
placement, hydrograph shapes) and documented where they are made.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import mesh as M
from .params import Params

CELL = 6.25  # m, cell size of the 2,048,000-cell reach


@dataclasses.dataclass
class Scenario:
    mesh: M.Mesh
    params: Params
    state: np.ndarray          # initial (nelt, 3)
    channel: np.ndarray        # bool per cell
    zone: np.ndarray | None = None        # zone id per cell
    n_zones: np.ndarray | None = None     # n per zone
    gauges: np.ndarray | None = None      # gauge cells


def _dims(ncells: int, aspect: float = 10.0):
    ny = max(4, int(round(np.sqrt(ncells / (2.0 * aspect)))))
    nx = max(4, int(round(ncells / (2.0 * ny))))
    return nx, ny


def reach(ncells: int, seed: int = 1, q_peak: float = 3000.0, t_ramp: float = 3600.0, iflux: int = 3,
          fricimplic: int = 1, hydrograph=None, length: float | None = None) -> Scenario:
    """River reach of about ncells triangles (2 per quad): channel 1/10 of
    the width, 3 m deep, bed slope 1/2000 + 0.5 m bars (300 m) + a seeded
    smooth field (0.3 m, 200 m), n 0.03 channel / 0.06 floodplain, inflow
    hydrograph (default: ramp 0 -> q_peak over t_ramp), fixed outlet level,
    dry threshold 1e-3 m, WAF + superbee (or HLLC when iflux = 1).  By
    default the cells are 6.25 m and the reach grows with ncells; with
    `length` the reach is length x length/10 and the cells scale instead."""
    nx, ny = _dims(ncells)
    if length is None:
        length, width = nx * CELL, ny * CELL
    else:
        width = length / 10.0
    r = M.river_mesh(nx=nx, ny=ny, length=length, width=width, channel_width=width / 10.0, q_peak=q_peak,
                     t_ramp=t_ramp, seed=seed)
    p = Params(iflux=iflux, ilimiteur=1, tolisec=1e-3, cfl=0.9, ts=1e300, inlet="inflow", outlet="fixedheight",
               friction=1, fricimplic=fricimplic, nombre_entrees=1, nombre_sorties=1)
    p.h_sortie[0] = r.h_out
    p.hydrographs = {1: hydrograph if hydrograph is not None else ([0.0, t_ramp], [0.0, q_peak])}
    return Scenario(mesh=r.mesh, params=p, state=r.state, channel=r.channel)


def river_reach(nx: int = 3200, ny: int = 320, order: str = "rcm", seed: int = 1):
    """The river-reach configuration: 20 km x 2 km (scaled with nx, ny)
    meshed as nx x ny jittered quads split into triangles (2,048,000 at the
    default size), stored in RCM, random or generation ("natural") order;
    bed slope 1/2000 + 0.5 m bars (300 m) + a seeded smooth field (0.3 m,
    200 m), a 3 m deep 200 m channel in a dry floodplain, n 0.03 / 0.06,
    inflow ramped 0 -> 3000 m3/s over 1 h, fixed outlet level, dry
    threshold 1e-3 m, WAF + superbee + fricimplic 1.  Returns (mesh,
    params, initial state, perm) with cell k the generated cell perm[k]."""
    r = M.river_mesh(nx=nx, ny=ny, length=20000.0 * nx / 3200, width=2000.0 * ny / 320, seed=seed)
    m, U0 = r.mesh, r.state
    if order == "rcm":
        perm = M.rcm_order(m)
    elif order == "random":
        perm = M.random_order(m, seed)
    elif order == "natural":
        perm = np.arange(m.nelt)
    else:
        raise ValueError(order)
    if order != "natural":
        m, U0 = m.permuted(perm), U0[perm]
    p = Params(iflux=3, ilimiteur=1, tolisec=1e-3, cfl=0.9, ts=1e300, inlet="inflow",
               outlet="fixedheight", friction=1, fricimplic=1, nombre_entrees=1, nombre_sorties=1)
    p.h_sortie[0] = r.h_out
    p.hydrographs = {1: ([0.0, r.t_ramp], [0.0, r.q_peak])}
    return m, p, np.ascontiguousarray(U0), perm


def zones_of(m: M.Mesh, channel: np.ndarray, nbands: int = 4) -> np.ndarray:
    """8 zones by default: nbands reaches along the flow, channel and
    floodplain in each (zone = 2 * band + floodplain)."""
    xc, _ = m.centroids()
    band = np.minimum((xc / xc.max() * nbands).astype(np.int64), nbands - 1)
    return 2 * band + (~channel).astype(np.int64)


def ensemble(nmembers: int = 32, nmin: int = 50_000, nmax: int = 500_000, seed: int = 2024, iflux: int = 3):
    """Members of uneven size (uniform in [nmin, nmax] cells), each the
    20 km x 2 km reach (200 m channel) meshed at that size, peak inflow
    uniform in 500-4000 m3/s, n uniform in 0.02-0.08 in each of 8 zones."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nmembers):
        n = int(rng.uniform(nmin, nmax))
        q = float(rng.uniform(500.0, 4000.0))
        s = reach(n, seed=seed + 7 * k + 1, q_peak=q, iflux=iflux, length=20000.0)
        s.zone = zones_of(s.mesh, s.channel)
        s.n_zones = rng.uniform(0.02, 0.08, 8)
        s.mesh.manning = s.n_zones[s.zone].copy()
        out.append(s)
    return out


def calibration(ncells: int = 100_000, ngauges: int = 20, seed: int = 7, iflux: int = 3,
                flood: float = 0.5) -> Scenario:
    """Zonal calibration reach: 8 zones (zones_of: 4 reaches along the flow x
    channel/floodplain), hidden truth n uniform in 0.025-0.05 per zone
    (Scenario.n_zones).  Every zone is active from the first step: the
    reach starts flooded, its water surface `flood` m above the mean
    floodplain level and inclined like the bed (1/2000), at rest, so that
    the water drains down the reach over channel and floodplain alike (the
    outlet holds that level, the inflow hydrograph adds its wave); bed highs
    of the floodplain stay dry.  Gauges: the same number in each zone (the
    remainder in the channel zones, upstream first), at seeded cells at
    least 0.2 m deep at the start."""
    rng = np.random.default_rng(seed)
    s = reach(ncells, seed=seed, iflux=iflux)
    s.zone = zones_of(s.mesh, s.channel)
    s.n_zones = rng.uniform(0.025, 0.05, 8)
    s.mesh.manning = s.n_zones[s.zone].copy()
    _flood(s, flood)
    s.gauges = _stratified_gauges(rng, s.zone, s.state, ngauges)
    return s


def _flood(s: Scenario, flood: float) -> None:
    """Initial state at rest with the water surface `flood` m above the mean
    floodplain level, inclined like the bed (1/2000); the outlet holds that
    level."""
    m = s.mesh
    xc, _ = m.centroids()
    length = float(m.coords[:, 0].max())
    zc = m.coords[m.conn - 1, 2].mean(axis=1)
    s.state = np.zeros((m.nelt, 3))
    s.state[:, 0] = np.maximum(-xc / 2000.0 + flood - zc, 0.0)
    s.params.h_sortie[0] = -length / 2000.0 + flood


def _stratified_gauges(rng, zone: np.ndarray, state: np.ndarray, ngauges: int) -> np.ndarray:
    """ngauges cells spread evenly over the 8 zones (the remainder in the
    channel zones, upstream first), each at least 0.2 m deep at the start."""
    per = np.full(8, ngauges // 8)
    per[[0, 2, 4, 6][: ngauges % 8]] += 1
    g = []
    for z in range(8):
        cand = np.flatnonzero((zone == z) & (state[:, 0] >= 0.2))
        g.append(rng.choice(cand, size=min(per[z], len(cand)), replace=False))
    return np.sort(np.concatenate(g)).astype(np.int32)


def calibration_batch(nscen: int = 64, ncells: int = 100_000, ngauges: int = 20, seed: int = 64,
                      iflux: int = 1):
    """nscen flood events on the 100k-cell reach (one mesh, 8 zones as in
    calibration(), the river configuration's start: water in the channel,
    the floodplain dry): scenario k has its own peak inflow (uniform in
    500-4000 m3/s over the 1 h ramp), hidden true n per zone (uniform in
    0.025-0.05) and ngauges seeded channel gauges in the downstream three
    quarters."""
    rng = np.random.default_rng(seed)
    base = reach(ncells, seed=seed, iflux=iflux)
    zone = zones_of(base.mesh, base.channel)
    xc, _ = base.mesh.centroids()
    cand = np.flatnonzero(base.channel & (xc > 0.25 * xc.max()))
    out = []
    for k in range(nscen):
        q = float(rng.uniform(500.0, 4000.0))
        p = dataclasses.replace(base.params)
        p.hydrographs = {1: ([0.0, 3600.0], [0.0, q])}
        nz = rng.uniform(0.025, 0.05, 8)
        m = dataclasses.replace(base.mesh, manning=nz[zone].copy())
        g = np.sort(rng.choice(cand, size=ngauges, replace=False)).astype(np.int32)
        out.append(Scenario(mesh=m, params=p, state=base.state.copy(), channel=base.channel, zone=zone,
                            n_zones=nz, gauges=g))
    return out


def field_ensemble(nmembers: int = 16, ncells: int = 1_000_000, ngauges: int = 50, seed: int = 99,
                   iflux: int = 3):
    """Members on the same 20 km x 2 km reach with a per-cell log-normal Manning field
    (median 0.04, sigma_log 0.3, 500 m correlation) and their own inflow
    hydrograph (rise to a peak uniform in 1500-4000 m3/s over 20-40 min, then
    recession), so that the floodplain wets and dries; 50 gauges per member
    at seeded channel cells."""
    rng = np.random.default_rng(seed)
    base = reach(ncells, seed=seed, iflux=iflux, length=20000.0)
    m0 = base.mesh
    xc, yc = m0.centroids()
    ext = (0.0, 0.0, float(m0.coords[:, 0].max()), float(m0.coords[:, 1].max()))
    cand = np.flatnonzero(base.channel & (xc > 0.1 * xc.max()))
    out = []
    for k in range(nmembers):
        field = M.smooth_random_field(xc, yc, 0.3, 500.0, seed + 100 + k, ext)
        n = 0.04 * np.exp(field)
        q = float(rng.uniform(1500.0, 4000.0))
        t1 = float(rng.uniform(1200.0, 2400.0))
        hydro = ([0.0, t1, 3.0 * t1], [0.0, q, 0.2 * q])
        p = dataclasses.replace(base.params)
        p.hydrographs = {1: hydro}
        m = dataclasses.replace(m0, manning=n)
        g = np.sort(rng.choice(cand, size=ngauges, replace=False)).astype(np.int32)
        out.append(Scenario(mesh=m, params=p, state=base.state.copy(), channel=base.channel, gauges=g))
    return out
