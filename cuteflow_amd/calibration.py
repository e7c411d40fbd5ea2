"""Manning calibration on the GPU adjoint (the reference's Calibration_manning/
workflow estimates zonal n by repeated forward runs).

- zonal_lbfgs(): L-BFGS on zonal n through GaugeLevels (torch.optim.LBFGS);
- zonal_fd_gradient(): central differences of the same loss on each zone,
  with the recorded dt (the map the adjoint differentiates), for checks;
- extent_loss(): a differentiable flood-extent term from the per-cell
  maximum depth (a smooth step of hmax weighted by cell area), for the
  flood-extent misfit of the per-cell field calibration.
"""
from __future__ import annotations

import numpy as np

from .adjoint import AdjointRun, GaugeLevels, zone_gather, zone_map


def gauge_misfit(eta, obs):
    return float(np.sum((np.asarray(eta) - np.asarray(obs)) ** 2))


def zonal_fd_gradient(run: AdjointRun, Z: np.ndarray, n_zones: np.ndarray, obs, eps: float = 1e-6,
                      count_flips: bool = False):
    """Central differences (relative step eps) of the gauge misfit with
    respect to each zone's n, at the recorded dt: the misfit at a perturbed
    n is taken from a tangent-linear run's values (the recorded run replayed
    with its dt and tc).  Single-instance runs.  Restores n afterwards.

    count_flips: also returns, per zone, the numbers of (step, cell) wet/dry
    clamp decisions and of (step, cell) updates with any branch taken
    differently by the two perturbed runs than by the run at n_zones (a
    finite difference across such a kink of the frozen-branch map differs
    from the derivative): (g, clamp_flips, branch_flips)."""
    obs = np.asarray(obs).reshape(-1)
    g = np.zeros(len(n_zones))
    flips = np.zeros(len(n_zones), np.int64)
    bflips = np.zeros(len(n_zones), np.int64)
    if count_flips:
        run.set_manning([Z @ np.asarray(n_zones, dtype=np.float64)])
        run.set_mask_mode(3)
        run.tlm()
        run.set_mask_mode(4)
        run.flips(0, reset=True)
        run.flips(1, reset=True)
    for z in range(len(n_zones)):
        vals = []
        for sgn in (1.0, -1.0):
            nz = np.array(n_zones, dtype=np.float64)
            nz[z] *= 1.0 + sgn * eps
            run.set_manning([Z @ nz])
            vals.append(gauge_misfit(run.tlm()[0]["eta"].reshape(-1), obs))
        g[z] = (vals[0] - vals[1]) / (2.0 * eps * n_zones[z])
        if count_flips:
            flips[z] = int(np.sum(run.flips(0, reset=True)[0], dtype=np.int64))
            bflips[z] = int(np.sum(run.flips(1, reset=True)[0], dtype=np.int64))
    if count_flips:
        run.set_mask_mode(0)
    run.set_manning([Z @ np.asarray(n_zones, dtype=np.float64)])
    return (g, flips, bflips) if count_flips else g


def zonal_tlm_gradient(run: AdjointRun, Z: np.ndarray, obs):
    """Directional derivatives of the gauge misfit along each zone's
    indicator (tangent-linear runs at the current n): the gradient by
    forward-mode differentiation, one run per zone."""
    obs = np.asarray(obs).reshape(-1)
    g = np.zeros(Z.shape[1])
    for z in range(Z.shape[1]):
        t = run.tlm([np.ascontiguousarray(Z[:, z])])[0]
        g[z] = np.sum(2.0 * (t["eta"].reshape(-1) - obs) * t["d_eta"].reshape(-1))
    return g


def zonal_lbfgs(run: AdjointRun, zone_of_cell, n_start, obs, nsteps: int, U0, iters: int = 20, device=None):
    """Calibrates zonal n against gauge observations obs (nsamp, ngauges) of
    a single-instance run: L-BFGS (strong Wolfe line search) on the misfit
    relative to its starting value.  Returns (n_zones, history of relative
    misfits, number of gradient evaluations)."""
    import torch

    dev = device or torch.device("cuda")
    nz = len(n_start)
    Z = zone_map(zone_of_cell, nz, device=dev)
    zidx = torch.as_tensor(np.asarray(zone_of_cell, np.int64), device=dev)
    obs_t = torch.as_tensor(np.asarray(obs).reshape(-1), device=dev)
    n_z = torch.tensor(np.asarray(n_start, dtype=np.float64), device=dev, requires_grad=True)
    if dev.type == "cuda":
        U0 = torch.as_tensor(np.asarray(U0, np.float64), device=dev)   # resident for every evaluation
    with torch.no_grad():
        l0 = ((GaugeLevels.apply(zone_gather(n_z, zidx, Z), run, nsteps, [U0]) - obs_t) ** 2).sum().item()
    opt = torch.optim.LBFGS([n_z], lr=1.0, max_iter=iters, history_size=10, tolerance_grad=1e-14,
                            tolerance_change=1e-16, line_search_fn="strong_wolfe")
    hist = []

    def closure():
        opt.zero_grad()
        eta = GaugeLevels.apply(zone_gather(n_z, zidx, Z), run, nsteps, [U0])
        loss = ((eta - obs_t) ** 2).sum() / l0
        loss.backward()
        hist.append(loss.item())
        return loss

    opt.step(closure)
    return n_z.detach().cpu().numpy(), hist, len(hist)


def soft_extent(hmax, surf, h0: float = 0.05, width: float = 0.02):
    """Flooded area sum(surf * sigmoid((hmax - h0) / width)) and its
    derivative with respect to hmax (per cell).  hmax and surf: numpy
    arrays, or torch tensors (the derivative is then a tensor on their
    device)."""
    if hasattr(hmax, "data_ptr"):
        import torch

        z = ((hmax - h0) / width).clamp(-60.0, 60.0)
        sg = 1.0 / (1.0 + torch.exp(-z))
        area = float(torch.sum(surf * sg))
        return area, surf * sg * (1.0 - sg) / width
    z = (np.asarray(hmax) - h0) / width
    sg = 1.0 / (1.0 + np.exp(-np.clip(z, -60.0, 60.0)))
    area = float(np.sum(surf * sg))
    return area, surf * sg * (1.0 - sg) / width


def extent_loss(hmax, surf, area_obs, scale, h0: float = 0.05, width: float = 0.02):
    """((A - A_obs) / scale)^2 of the soft flooded area A and its cotangent
    with respect to hmax."""
    a, da = soft_extent(hmax, surf, h0, width)
    r = (a - area_obs) / scale
    return r * r, (2.0 * r / scale) * da
