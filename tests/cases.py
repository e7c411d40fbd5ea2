"""Seeded test cases shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

from cuteflow_amd import mesh as M
from cuteflow_amd.params import Params


def dam_case(n=20, seed=0):
    """Circular dam break on an n x n jittered square, HLLC, walls, flat bed."""
    m = M.dam_break_mesh(n, size=200.0, jitter=0.2, seed=seed)
    p = Params(iflux=1, tolisec=1e-5, cfl=0.9, ts=1e300, inlet="transm", outlet="transm", friction=0)
    U0 = M.circular_dam_state(m)
    return m, p, U0


def river_case(nx=60, ny=12, seed=3, iflux=3, fricimplic=1, ilimiteur=1, friction=1,
               q_peak=600.0, t_ramp=50.0, outlet="fixedheight"):
    """Scaled-down river reach with floodplain, inflow hydrograph and a fixed
    outlet level; WAF with semi-implicit friction by default."""
    length, width = 600.0 * nx / 60, 200.0 * ny / 12
    r = M.river_mesh(nx=nx, ny=ny, length=length, width=width, channel_width=width / 3,
                     depth=3.0, bar_len=120.0, field_corr=60.0, q_peak=q_peak, t_ramp=t_ramp,
                     seed=seed)
    p = Params(iflux=iflux, ilimiteur=ilimiteur, tolisec=1e-3, cfl=0.9, ts=1e300, inlet="inflow",
               outlet=outlet, friction=friction, fricimplic=fricimplic, nombre_entrees=1,
               nombre_sorties=1)
    p.h_sortie[0] = r.h_out
    p.hydrographs = {1: ([0.0, t_ramp], [0.0, q_peak])}
    return r.mesh, p, r.state


def lake_case(n=12, seed=5, iflux=1):
    """Lake at rest over a rough bed with dry islands: must stay at rest."""
    m = M.dam_break_mesh(n, size=100.0, jitter=0.2, seed=seed)
    rng = np.random.default_rng(seed)
    m.coords[:, 2] = rng.uniform(-2.0, 0.5, m.nnt)
    p = Params(iflux=iflux, tolisec=1e-5, cfl=0.9, ts=1e300, inlet="transm", outlet="transm",
               friction=1, fricimplic=1)
    return m, p
