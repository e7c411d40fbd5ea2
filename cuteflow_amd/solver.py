"""Host-side mirror of the reference time loop around the C ABI.

`Solver` owns one cf_handle (one subdomain).  `solve_finite_volumes` is the
reference driver loop (finite_volumes_method.cuf:4-258): initial dt, then
steps until tc >= ts - tol, with the per-step work done on the device in
batches of steps (no host synchronisation inside a batch).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .mesh import Mesh
from .params import Params


class Solver:
    def __init__(self, mesh: Mesh, params: Params, boundary: np.ndarray | None = None,
                 ts: float | None = None, options: int = 0):
        """options: CF_OPT_* bits (device layout only; results are identical)."""
        lib = _lib.load()
        self.lib = lib
        self.mesh = mesh
        self.params = params
        self.boundary = mesh.boundary_table(device=True) if boundary is None else np.ascontiguousarray(boundary, np.int32)
        self.long_entree = mesh.inlet_lengths(self.boundary, params.nombre_entrees)
        self.cfp = params.to_cf(self.long_entree, ts=ts)
        self.cfp.options = int(options)
        self._conn = np.ascontiguousarray(mesh.conn.reshape(-1), np.int32)
        self._coords = np.ascontiguousarray(mesh.coords.reshape(-1), np.float64)
        self._mann = np.ascontiguousarray(mesh.manning, np.float64)
        cm = _lib.CfMesh()
        cm.nelt = mesh.nelt
        cm.nnt = mesh.nnt
        cm.conn = _lib.iptr(self._conn)
        cm.coords = _lib.dptr(self._coords)
        cm.boundary = _lib.iptr(self.boundary)
        cm.manning = _lib.dptr(self._mann)
        h = C.c_void_p()
        _lib.check(lib.cf_create(C.byref(cm), C.byref(self.cfp), C.byref(h)), "cf_create")
        self.h = h
        self.nelt = mesh.nelt

    # -- state ---------------------------------------------------------------
    def set_state(self, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, np.float64)
        assert U.shape == (self.nelt, 3)
        _lib.check(self.lib.cf_set_state(self.h, _lib.dptr(U)), "cf_set_state")

    def get_state(self) -> np.ndarray:
        U = np.zeros((self.nelt, 3))
        _lib.check(self.lib.cf_get_state(self.h, _lib.dptr(U)), "cf_get_state")
        return U

    def set_state_device(self, ptr: int) -> None:
        _lib.check(self.lib.cf_set_state_device(self.h, C.c_void_p(ptr)), "cf_set_state_device")

    def get_state_device(self, ptr: int) -> None:
        _lib.check(self.lib.cf_get_state_device(self.h, C.c_void_p(ptr)), "cf_get_state_device")

    @property
    def dt(self) -> float:
        v = C.c_double()
        _lib.check(self.lib.cf_initial_dt(self.h, C.byref(v)), "cf_initial_dt")
        return v.value

    @property
    def tc(self) -> float:
        t = C.c_double()
        _lib.check(self.lib.cf_get_time(self.h, C.byref(t), None), "cf_get_time")
        return t.value

    def geometry(self):
        n = self.nelt
        surf, zm, cot = np.zeros(n), np.zeros(n), np.zeros(n)
        cmin = C.c_double()
        _lib.check(self.lib.cf_get_geometry(self.h, _lib.dptr(surf), _lib.dptr(zm), _lib.dptr(cot),
                                            C.byref(cmin)), "cf_get_geometry")
        return {"surf": surf, "zm": zm, "cotemin_arr": cot, "cotemin": cmin.value}

    # -- stepping --------------------------------------------------------------
    def step(self, nsteps: int, raise_on_error: bool = True):
        """Run up to nsteps steps; returns (dt_hist[:done+1], done, status)."""
        hist = np.zeros(nsteps + 1)
        done = C.c_int()
        st = self.lib.cf_step(self.h, int(nsteps), _lib.dptr(hist), C.byref(done))
        self.last_done = done.value
        if st != _lib.CF_OK and (raise_on_error or st != _lib.CF_ERR_DT):
            raise _lib.CuteflowError(st, "cf_step")
        return hist[: done.value + 1], done.value, st

    def prepare(self, nsteps: int) -> None:
        """One-time work of runs of nsteps steps done now (history sized, the
        run's graph captured): the next step(nsteps) launches only."""
        _lib.check(self.lib.cf_step_prepare(self.h, int(nsteps)), "cf_step_prepare")

    def set_discharge(self, on: bool = True) -> None:
        """Per-step inflow/outflow discharge of the following runs (always on
        when tol_reg_perm > 0: the steady-state stop uses it)."""
        _lib.check(self.lib.cf_set_discharge(self.h, int(bool(on))), "cf_set_discharge")

    def discharge(self):
        """(q_in, q_out, steady) of the last step(): per-step inflow and
        outflow discharges and whether the run stopped at steady state."""
        n = getattr(self, "last_done", 0)
        q_in, q_out, steady = np.zeros(n), np.zeros(n), C.c_int()
        _lib.check(self.lib.cf_get_discharge(self.h, _lib.dptr(q_in), _lib.dptr(q_out), C.byref(steady)),
                   "cf_get_discharge")
        return q_in, q_out, bool(steady.value)

    def last_run_stats(self) -> dict:
        """Tiles that ran in the last step() (dry tiles with dry neighbours are
        skipped) out of tiles x steps."""
        run, tot, cells = C.c_longlong(), C.c_longlong(), C.c_int()
        _lib.check(self.lib.cf_last_run_stats(self.h, C.byref(run), C.byref(tot), C.byref(cells)),
                   "cf_last_run_stats")
        return {"tiles_run": run.value, "tiles_total": tot.value, "tile_cells": cells.value,
                "active_fraction": run.value / tot.value if tot.value else 0.0}

    def set_manning(self, n: np.ndarray) -> None:
        n = np.ascontiguousarray(n, np.float64)
        assert n.shape == (self.nelt,)
        _lib.check(self.lib.cf_set_manning(self.h, _lib.dptr(n)), "cf_set_manning")

    def set_inflow(self, q) -> None:
        q = np.ascontiguousarray(q, np.float64)
        _lib.check(self.lib.cf_set_inflow(self.h, _lib.dptr(q), len(q)), "cf_set_inflow")

    def set_hydrograph(self, inlet: int, t, q) -> None:
        t = np.ascontiguousarray(t, np.float64)
        q = np.ascontiguousarray(q, np.float64)
        _lib.check(self.lib.cf_set_hydrograph(self.h, inlet, len(t), _lib.dptr(t), _lib.dptr(q)),
                   "cf_set_hydrograph")

    @property
    def stream(self) -> int:
        return self.lib.cf_get_stream(self.h) or 0

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.cf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _batch_structs(meshes, params, ts, options):
    """cf_mesh / cf_params arrays of a batch (boundary tables and inlet
    lengths derived from each mesh), the numpy buffers they point into, and
    the boundary tables."""
    n = len(meshes)
    cms = (_lib.CfMesh * n)()
    cps = (_lib.CfParams * n)()
    keep, boundaries = [], []
    shared = {}  # meshes sharing their arrays (scenarios of one reach) share the tables
    for b, (m, p) in enumerate(zip(meshes, params)):
        key = (id(m.conn), id(m.coords), id(m.ndinput), id(m.numero_ndinput), id(m.ndoutput),
               id(m.numero_ndoutput), id(m.ndwall), p.nombre_entrees)
        if key not in shared:
            bt = m.boundary_table(device=True)
            shared[key] = (bt, m.inlet_lengths(bt, p.nombre_entrees), np.ascontiguousarray(m.conn.reshape(-1), np.int32),
                           np.ascontiguousarray(m.coords.reshape(-1), np.float64))
        bt, le, conn, coords = shared[key]
        mann = np.ascontiguousarray(m.manning, np.float64)
        keep += [bt, conn, coords, mann]
        boundaries.append(bt)
        cms[b].nelt = m.nelt
        cms[b].nnt = m.nnt
        cms[b].conn = _lib.iptr(conn)
        cms[b].coords = _lib.dptr(coords)
        cms[b].boundary = _lib.iptr(bt)
        cms[b].manning = _lib.dptr(mann)
        cps[b] = p.to_cf(le, ts=ts)
        cps[b].options = int(options)
    return cms, cps, keep, boundaries


def tune_candidate_options(k: int, base: int = 0) -> int:
    """Options bits of autotuner candidate k (cf_batch_autotune's order)."""
    o = base & ~(_lib.CF_OPT_TILE_128 | _lib.CF_OPT_TILE_64 | _lib.CF_OPT_NO_TILE_SKIP | _lib.CF_OPT_XCD_CONTIG)
    o |= (0, _lib.CF_OPT_TILE_128, _lib.CF_OPT_TILE_64)[k // 4]
    if k & 2:
        o |= _lib.CF_OPT_NO_TILE_SKIP
    if k & 1:
        o |= _lib.CF_OPT_XCD_CONTIG
    return o


def autotune(meshes, params, states, nsteps: int = 100, ts=None, options: int = 0):
    """Times the batch under every schedule candidate (cells per tile 256,
    128, 64; dry-tile skipping on/off; XCD-contiguous tiles off/on) from the
    given initial states and returns (best_options, table), table being a
    list of (options, milliseconds or None) per candidate.  The best is the
    fastest, except that a candidate skipping dry tiles within 3 % of a
    faster one that does not is preferred (the skip also shortens the
    derivative runs' replay and reverse sweep).  Pass
    best_options to BatchSolver(options=...).  Results of a run do not depend
    on the choice, only its speed."""
    lib = _lib.load()
    if isinstance(params, Params):
        params = [params] * len(meshes)
    if len(params) != len(meshes) or len(states) != len(meshes) or not meshes:
        raise ValueError("one Params and one state per mesh")
    cms, cps, keep, _ = _batch_structs(meshes, params, ts, options)
    st = []
    for m, U in zip(meshes, states):
        U = np.ascontiguousarray(U, np.float64)
        if U.shape != (m.nelt, 3):
            raise ValueError("state shape must be (nelt, 3)")
        st.append(U)
    ptrs = (C.c_void_p * len(st))(*[U.ctypes.data for U in st])
    best = C.c_int()
    times = np.zeros(_lib.CF_TUNE_NCAND)
    _lib.check(lib.cf_batch_autotune(len(meshes), cms, cps, ptrs, int(nsteps), C.byref(best), _lib.dptr(times)),
               "cf_batch_autotune")
    table = [(tune_candidate_options(k, options), float(t) if t >= 0 else None) for k, t in enumerate(times)]
    return best.value, table


class BatchSolver:
    """Many independent subdomains advanced together (cf_batch_* entry
    points): one launch per step covers every instance; each instance keeps
    its own parameters, state, dt, time and status.  All instances share the
    scheme (iflux), friction mode and options."""

    def __init__(self, meshes, params, ts=None, options: int = 0):
        lib = _lib.load()
        self.lib = lib
        if isinstance(params, Params):
            params = [params] * len(meshes)
        if len(params) != len(meshes) or not meshes:
            raise ValueError("one Params per mesh")
        self.meshes = list(meshes)
        self.params = list(params)
        n = len(meshes)
        cms, cps, keep, self.boundaries = _batch_structs(meshes, params, ts, options)
        h = C.c_void_p()
        _lib.check(lib.cf_batch_create(n, cms, cps, C.byref(h)), "cf_batch_create")
        self.h = h
        self.ninst = n
        self.nelt = [m.nelt for m in meshes]

    def set_state(self, b: int, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, np.float64)
        assert U.shape == (self.nelt[b], 3)
        _lib.check(self.lib.cf_batch_set_state(self.h, b, _lib.dptr(U)), "cf_batch_set_state")

    def get_state(self, b: int) -> np.ndarray:
        U = np.zeros((self.nelt[b], 3))
        _lib.check(self.lib.cf_batch_get_state(self.h, b, _lib.dptr(U)), "cf_batch_get_state")
        return U

    def set_state_device(self, b: int, ptr: int) -> None:
        _lib.check(self.lib.cf_batch_set_state_device(self.h, b, C.c_void_p(ptr)), "cf_batch_set_state_device")

    def get_state_device(self, b: int, ptr: int) -> None:
        _lib.check(self.lib.cf_batch_get_state_device(self.h, b, C.c_void_p(ptr)), "cf_batch_get_state_device")

    def time(self, b: int):
        """(tc, dt) of instance b."""
        t, d = C.c_double(), C.c_double()
        _lib.check(self.lib.cf_batch_get_time(self.h, b, C.byref(t), C.byref(d)), "cf_batch_get_time")
        return t.value, d.value

    def set_manning(self, b: int, n: np.ndarray) -> None:
        n = np.ascontiguousarray(n, np.float64)
        assert n.shape == (self.nelt[b],)
        _lib.check(self.lib.cf_batch_set_manning(self.h, b, _lib.dptr(n)), "cf_batch_set_manning")

    def set_hydrograph(self, b: int, inlet: int, t, q) -> None:
        t = np.ascontiguousarray(t, np.float64)
        q = np.ascontiguousarray(q, np.float64)
        _lib.check(self.lib.cf_batch_set_hydrograph(self.h, b, inlet, len(t), _lib.dptr(t), _lib.dptr(q)),
                   "cf_batch_set_hydrograph")

    def set_inflow(self, b: int, q) -> None:
        q = np.ascontiguousarray(q, np.float64)
        _lib.check(self.lib.cf_batch_set_inflow(self.h, b, _lib.dptr(q), len(q)), "cf_batch_set_inflow")

    def geometry(self, b: int):
        n = self.nelt[b]
        surf, zm, cot = np.zeros(n), np.zeros(n), np.zeros(n)
        cmin = C.c_double()
        _lib.check(self.lib.cf_batch_get_geometry(self.h, b, _lib.dptr(surf), _lib.dptr(zm), _lib.dptr(cot),
                                                  C.byref(cmin)), "cf_batch_get_geometry")
        return {"surf": surf, "zm": zm, "cotemin_arr": cot, "cotemin": cmin.value}

    def step(self, nsteps: int, raise_on_error: bool = False):
        """Run up to nsteps steps of every running instance.  Returns
        (dt_hist list per instance [:done+1], done array, status array)."""
        B = self.ninst
        hist = np.zeros((B, nsteps + 1))
        done = np.zeros(B, np.int32)
        status = np.zeros(B, np.int32)
        st = self.lib.cf_batch_step(self.h, int(nsteps), _lib.dptr(hist), _lib.iptr(done), _lib.iptr(status))
        self.last_done = done.copy()
        if st != _lib.CF_OK and (raise_on_error or st != _lib.CF_ERR_DT):
            raise _lib.CuteflowError(st, "cf_batch_step")
        return [hist[b, : done[b] + 1] for b in range(B)], done, status

    def prepare(self, nsteps: int) -> None:
        """One-time work of runs of nsteps steps done now (see Solver.prepare)."""
        _lib.check(self.lib.cf_step_prepare(self.h, int(nsteps)), "cf_step_prepare")

    def set_discharge(self, on: bool = True) -> None:
        _lib.check(self.lib.cf_set_discharge(self.h, int(bool(on))), "cf_set_discharge")

    def discharge(self, b: int):
        """(q_in, q_out, steady) of instance b in the last step()."""
        n = int(self.last_done[b]) if hasattr(self, "last_done") else 0
        q_in, q_out, steady = np.zeros(n), np.zeros(n), C.c_int()
        _lib.check(self.lib.cf_batch_get_discharge(self.h, b, _lib.dptr(q_in), _lib.dptr(q_out), C.byref(steady)),
                   "cf_batch_get_discharge")
        return q_in, q_out, bool(steady.value)

    def last_run_stats(self) -> dict:
        run, tot, cells = C.c_longlong(), C.c_longlong(), C.c_int()
        _lib.check(self.lib.cf_last_run_stats(self.h, C.byref(run), C.byref(tot), C.byref(cells)),
                   "cf_last_run_stats")
        return {"tiles_run": run.value, "tiles_total": tot.value, "tile_cells": cells.value,
                "active_fraction": run.value / tot.value if tot.value else 0.0}

    @property
    def stream(self) -> int:
        return self.lib.cf_get_stream(self.h) or 0

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.cf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def initial_state(params: Params, mesh: Mesh, zm: np.ndarray) -> np.ndarray:
    """set_initial_condition (initial_condition.cuf:4-80): dam line or plane
    initial condition from the centroid and the cell bed level zm."""
    xc, yc = mesh.centroids()
    U = np.zeros((mesh.nelt, 3))
    if params.eqlin_barrage == 1:
        alpha = (params.y1eqbar - params.y2eqbar) / (params.x1eqbar - params.x2eqbar)
        beta = params.y1eqbar - alpha * params.x1eqbar
        up = (alpha * xc - yc) + beta >= 0.0
        h = np.where(up, params.h_amont - zm, params.h_aval - zm)
        u = np.where(up, params.u_amont, params.u_aval)
        v = np.where(up, params.v_amont, params.v_aval)
        wet = h > params.tolisec
        U[:, 0] = np.where(wet, h, 0.0)
        U[:, 1] = np.where(wet, h * u, 0.0)
        U[:, 2] = np.where(wet, h * v, 0.0)
    if params.plan == 1:
        h = ((((-params.xplan) * (params.xpoint - xc)) - params.yplan * (params.ypoint - yc)) / params.zplan
             + params.zpoint) - zm
        wet = h > params.tolisec
        U[:, 0] = np.where(wet, h, 0.0)
        U[:, 1] = np.where(wet, h * params.u_aval, 0.0)
        U[:, 2] = np.where(wet, h * params.v_aval, 0.0)
    return U


def solve_finite_volumes(solver: Solver, batch: int = 1000, max_steps: int | None = None,
                         on_batch=None):
    """The reference loop (finite_volumes_method.cuf:113-242) until tc >= ts - tol
    (or max_steps): returns (dt_hist, steps, tc)."""
    hists, total = [], 0
    while True:
        n = batch if max_steps is None else min(batch, max_steps - total)
        if n <= 0:
            break
        hist, done, _ = solver.step(n)
        hists.append(hist[:done])
        total += done
        if on_batch is not None:
            on_batch(solver, total)
        if done < n:
            break
    dt_hist = np.concatenate(hists) if hists else np.zeros(0)
    return dt_hist, total, solver.tc
