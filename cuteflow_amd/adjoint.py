"""Gradients of gauge water levels with respect to Manning n (discrete
adjoint of the HIP time step) for the calibration workflow.

The reference calibrates n zone by zone with repeated forward runs driven by
scripts (Calibration_manning/); here the derivative of the discrete step is
exact (wet/dry switches, wave-speed cases, limiter regions and dt frozen as
taken in the forward), computed by a reverse sweep on the GPU with
checkpoint/recompute, and exposed to PyTorch as an autograd Function:

    run = AdjointRun(solver, gauges, sample_every=10, checkpoint_every=100)
    n_cells = zone_map(zone_of_cell, 8) @ n_zones          # differentiable
    eta = GaugeLevels.apply(n_cells, run, 2000, [U0])       # (nsamp, ngauges)
    loss = ((eta - eta_obs) ** 2).sum(); loss.backward()    # -> n_zones.grad

For a batch (BatchSolver) the Function takes the instances' per-cell n
concatenated and returns the instances' gauge series concatenated.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _ptr(a, sync: bool = True):
    """(pointer, keep-alive) for a per-cell array: numpy (host) or a
    contiguous float64 CUDA tensor (device; with sync, after the work queued
    on torch's current stream, which may still be writing it)."""
    if a is None:
        return None, None
    if hasattr(a, "data_ptr"):
        import torch

        t = a.detach().to(torch.float64).contiguous()
        if t.is_cuda:
            if sync:
                torch.cuda.current_stream().synchronize()
            return C.c_void_p(t.data_ptr()), t
        a = t.numpy()
    a = np.ascontiguousarray(a, np.float64)
    return a.ctypes.data_as(C.c_void_p), a


def _empty(n, device: bool):
    """A float64 array of n values: a CUDA tensor with device, else numpy."""
    if device:
        import torch

        return torch.empty(n, dtype=torch.float64, device="cuda")
    return np.zeros(n)


class AdjointRun:
    """Recorded forward runs and reverse sweeps on a Solver or BatchSolver.

    gauges: for a Solver one array of cell indices (caller's order), for a
    BatchSolver one array per instance.  freeze_limiters: derivatives with the
    WAF limiter values held at their forward values (CF_ADJ_FREEZE_LIMITERS);
    the exact derivative of the WAF step grows without bound over long runs
    through water at rest or uniform flow.
    """

    def __init__(self, solver, gauges, sample_every: int = 10, checkpoint_every: int = 100,
                 want_hmax: bool = False, freeze_limiters: bool = False):
        self.solver = solver
        self.lib = solver.lib
        self.h = solver.h
        self.batch = hasattr(solver, "ninst")
        self.ninst = solver.ninst if self.batch else 1
        self.nelt = list(solver.nelt) if self.batch else [solver.nelt]
        if not self.batch:
            gauges = [gauges]
        if len(gauges) != self.ninst:
            raise ValueError("one gauge list per instance")
        self.gauges = [np.ascontiguousarray(g, np.int32) for g in gauges]
        _lib.check(self.lib.cf_adj_set_flags(self.h, _lib.CF_ADJ_FREEZE_LIMITERS if freeze_limiters else 0),
                   "cf_adj_set_flags")
        for b, g in enumerate(self.gauges):
            _lib.check(self.lib.cf_adj_set_gauges(self.h, b, len(g), _lib.iptr(g)), "cf_adj_set_gauges")
        self.sample_every = int(sample_every)
        self.checkpoint_every = int(checkpoint_every)
        self.want_hmax = bool(want_hmax)
        self.nsteps = 0
        self.nsamp = 0
        self.records = 0          # recorded runs so far (the autograd Function checks it)

    # -- forward ---------------------------------------------------------------
    def set_manning(self, n_list) -> None:
        for b, n in enumerate(n_list):
            p, keep = _ptr(n)
            _lib.check(self.lib.cf_batch_set_manning(self.h, b, p), "cf_batch_set_manning")
            del keep

    def set_state(self, U_list) -> None:
        """Initial states (nelt, 3) per instance: numpy arrays, or float64
        CUDA tensors (copied on the device)."""
        for b, U in enumerate(U_list):
            if hasattr(U, "data_ptr") and U.is_cuda:
                import torch

                t = U.detach().to(torch.float64).contiguous()
                torch.cuda.current_stream().synchronize()
                _lib.check(self.lib.cf_batch_set_state_device(self.h, b, C.c_void_p(t.data_ptr())),
                           "cf_batch_set_state_device")
                continue
            U = np.ascontiguousarray(U, np.float64)
            _lib.check(self.lib.cf_batch_set_state(self.h, b, _lib.dptr(U)), "cf_batch_set_state")

    def forward(self, nsteps: int, U0=None, manning=None, device: bool = False):
        """Recorded run of nsteps steps (from U0 when given, else from the
        current state).  Returns (eta per instance (nsamp, ngauges), hmax per
        instance or None); with device, hmax as float64 CUDA tensors."""
        if manning is not None:
            self.set_manning(manning)
        if U0 is not None:
            self.set_state(U0)
        _lib.check(self.lib.cf_adj_forward(self.h, int(nsteps), self.sample_every, self.checkpoint_every,
                                           int(self.want_hmax)), "cf_adj_forward")
        self.nsteps = int(nsteps)
        self.records += 1
        self.nsamp = self.lib.cf_adj_num_samples(self.h)
        etas, hmaxs = [], []
        for b in range(self.ninst):
            eta = np.zeros((self.nsamp, len(self.gauges[b])))
            hm, hp = None, None
            if self.want_hmax:
                hm = _empty(self.nelt[b], device)
                hp = C.c_void_p(hm.data_ptr()) if device else hm.ctypes.data_as(C.c_void_p)
            _lib.check(self.lib.cf_adj_get_samples(self.h, b, _lib.dptr(eta), hp), "cf_adj_get_samples")
            etas.append(eta)
            hmaxs.append(hm)
        return etas, (hmaxs if self.want_hmax else None)

    # -- reverse ---------------------------------------------------------------
    def backward(self, g_eta, g_hmax=None, g_state=None, want_state: bool = True, device: bool = False,
                 out=None):
        """Cotangents of the outputs (per instance; None = 0; per-cell seeds
        may be CUDA tensors) -> gradients (per instance) with respect to
        per-cell n and the initial state (None unless want_state).  With
        device the n gradients are float64 CUDA tensors (written into the
        tensors of `out` when given)."""
        # per-cell seeds as pointers (CUDA tensors converted first, then one
        # synchronisation with torch's stream for all of them)
        gh_all = [_ptr(None if g_hmax is None else g_hmax[b], sync=False) for b in range(self.ninst)]
        gs_all = [_ptr(None if g_state is None else g_state[b], sync=False) for b in range(self.ninst)]
        if any(hasattr(k, "is_cuda") and k.is_cuda for _, k in gh_all + gs_all):
            import torch

            torch.cuda.current_stream().synchronize()
        for b in range(self.ninst):
            ge = None if g_eta is None or g_eta[b] is None else np.ascontiguousarray(g_eta[b], np.float64)
            (gh, kh), (gs, ks) = gh_all[b], gs_all[b]
            _lib.check(self.lib.cf_adj_set_seed(self.h, b, _lib.dptr(ge) if ge is not None else None, gh, gs),
                       "cf_adj_set_seed")
            del kh, ks
        _lib.check(self.lib.cf_adj_backward(self.h), "cf_adj_backward")
        gns, gus = [], []
        for b in range(self.ninst):
            gn = out[b] if out is not None else _empty(self.nelt[b], device)
            gu = np.zeros((self.nelt[b], 3)) if want_state else None
            pn = C.c_void_p(gn.data_ptr()) if device else gn.ctypes.data_as(C.c_void_p)
            _lib.check(self.lib.cf_adj_get_gradient(self.h, b, pn, gu.ctypes.data_as(C.c_void_p) if want_state
                                                    else None), "cf_adj_get_gradient")
            gns.append(gn)
            gus.append(gu)
        return gns, gus

    # -- tangent-linear --------------------------------------------------------
    def set_mask_mode(self, mode: int) -> None:
        """Branch masks of the next tangent-linear runs: 1 records the wet/dry
        clamp decision of each step and cell, 2 counts per cell the steps that
        differ from the recorded ones; 3 and 4 the same plus a signature of
        every branch of the cell's update; 0 off."""
        _lib.check(self.lib.cf_tlm_set_mask_mode(self.h, int(mode)), "cf_tlm_set_mask_mode")

    def flips(self, kind: int = 0, reset: bool = True):
        """Per instance, per cell: steps whose wet/dry decision (kind 0) or
        any branch (kind 1) differed from the recorded run in the counting
        runs (zeroed when reset)."""
        out = []
        for b in range(self.ninst):
            f = np.zeros(self.nelt[b], np.int32)
            _lib.check(self.lib.cf_tlm_get_flips(self.h, b, int(kind), _lib.iptr(f), int(bool(reset))),
                       "cf_tlm_get_flips")
            out.append(f)
        return out

    def tlm(self, dn=None, dU0=None):
        """Tangent-linear run over the recorded run along (dn, dU0) (per
        instance lists; None = 0).  Returns per instance a dict with eta and
        state (the run with the recorded dt replayed at the current n) and
        d_eta, d_hmax, d_state."""
        for b in range(self.ninst):
            pn, kn = _ptr(None if dn is None else dn[b])
            pu, ku = _ptr(None if dU0 is None else dU0[b])
            _lib.check(self.lib.cf_tlm_set_direction(self.h, b, pn, pu), "cf_tlm_set_direction")
            del kn, ku
        _lib.check(self.lib.cf_tlm_run(self.h), "cf_tlm_run")
        out = []
        for b in range(self.ninst):
            eta = np.zeros((self.nsamp, len(self.gauges[b])))
            d_eta = np.zeros((self.nsamp, len(self.gauges[b])))
            d_hmax = np.zeros(self.nelt[b]) if self.want_hmax else None
            d_state = np.zeros((self.nelt[b], 3))
            state = np.zeros((self.nelt[b], 3))
            _lib.check(self.lib.cf_tlm_get(self.h, b, _lib.dptr(eta), _lib.dptr(d_eta),
                                           d_hmax.ctypes.data_as(C.c_void_p) if d_hmax is not None else None,
                                           d_state.ctypes.data_as(C.c_void_p), state.ctypes.data_as(C.c_void_p)),
                       "cf_tlm_get")
            out.append({"eta": eta, "d_eta": d_eta, "d_hmax": d_hmax, "d_state": d_state, "state": state})
        return out


def _split(t, sizes):
    out, o = [], 0
    for s in sizes:
        out.append(t[o:o + s])
        o += s
    return out


class GaugeLevels:
    """torch.autograd.Function: per-cell Manning n (instances concatenated)
    -> gauge water levels h + zm (instances' (nsamp, ngauges) blocks
    flattened and concatenated), and the per-cell maximum depth over the
    samples when the run keeps it."""

    @staticmethod
    def apply(n_cells, run: AdjointRun, nsteps: int, U0):
        return _fn().apply(n_cells, run, nsteps, U0)


_FN = None


def _fn():
    global _FN
    if _FN is None:
        import torch

        class _GaugeLevels(torch.autograd.Function):
            @staticmethod
            def forward(ctx, n_cells, run, nsteps, U0):
                ns = _split(n_cells.detach(), run.nelt)
                on_dev = n_cells.is_cuda
                etas, hmaxs = run.forward(nsteps, U0=U0, manning=ns, device=on_dev)
                ctx.run = run
                ctx.record = run.records
                ctx.device = n_cells.device
                eta = torch.as_tensor(np.concatenate([e.reshape(-1) for e in etas]), device=n_cells.device)
                if run.want_hmax:
                    hm = torch.cat(hmaxs) if on_dev else torch.as_tensor(np.concatenate(hmaxs))
                    return eta, hm
                return eta

            @staticmethod
            def backward(ctx, g_eta, g_hmax=None):
                run = ctx.run
                if run.records != ctx.record:
                    raise RuntimeError("GaugeLevels.backward: the run was recorded again by a later forward "
                                       "(one AdjointRun holds one recorded run): call backward before the next "
                                       "forward on the same AdjointRun")
                sizes = [run.nsamp * len(g) for g in run.gauges]
                ge = [g.detach().cpu().numpy().reshape(run.nsamp, -1) if g.numel() else None
                      for g in _split(g_eta, sizes)]
                gh = None
                if run.want_hmax and g_hmax is not None:
                    gh = [g.detach() for g in _split(g_hmax, run.nelt)]
                if ctx.device.type == "cuda":
                    # the per-cell gradients are written into one CUDA tensor
                    g = torch.empty(sum(run.nelt), dtype=torch.float64, device=ctx.device)
                    run.backward(ge, gh, want_state=False, device=True, out=list(_split(g, run.nelt)))
                else:
                    gns, _ = run.backward(ge, gh, want_state=False)
                    g = torch.as_tensor(np.concatenate(gns))
                return g, None, None, None

        _FN = _GaugeLevels
    return _FN


def zone_gather(n_zones, zone_idx, Z):
    """Per-cell n from zonal n, differentiable: the forward is the gather
    n_zones[zone_idx] (exactly Z @ n_zones, without the dense product), the
    backward Z^T g as a column sum of Z * g, a reduction with a fixed order
    (an index-add would sum in atomic order).  Matrix-vector products for it
    are slow at this shape: Z.t() @ g runs a kernel with only nzones rows
    (1.9 ms for 100k cells), g @ Z a GEMM (10.7 ms).
    zone_idx: int64 tensor of each cell's zone; Z: zone_map of the same
    zones."""
    import torch

    class _ZoneGather(torch.autograd.Function):
        @staticmethod
        def forward(ctx, n):
            return n.index_select(0, zone_idx)

        @staticmethod
        def backward(ctx, g):
            return (Z * g.unsqueeze(1)).sum(dim=0)

    return _ZoneGather.apply(n_zones)


def zone_map(zone_of_cell, nzones: int, device=None):
    """Dense one-hot (ncells, nzones) float64 matrix: n_cells = Z @ n_zones,
    whose gradient Z^T g sums the per-cell gradients of each zone with a
    fixed reduction order."""
    import torch

    z = np.asarray(zone_of_cell, np.int64)
    Z = np.zeros((len(z), nzones))
    Z[np.arange(len(z)), z] = 1.0
    return torch.as_tensor(Z, device=device)
