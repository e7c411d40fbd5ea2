"""Discrete adjoint and tangent-linear model on the GPU (through the C ABI).

- the GPU tangent-linear run reproduces the host build of the same cell body
  bit for bit (values and tangents), and its values are the recorded
  forward's bits;
- the GPU reverse sweep equals the host reverse sweep (same contributions,
  summed in another order: 1e-12 relative);
- dot-product test on the GPU: <seeds, TLM outputs> = <gradients,
  direction> with gauge, max-depth and final-state seeds, over several
  checkpoint segments;
- gradient of a gauge misfit with respect to zonal Manning n (through the
  torch.autograd.Function) against the tangent-linear directional
  derivatives and central differences of the forward with the recorded dt;
- a few L-BFGS iterations on zonal n reduce the misfit;
- each instance of a batch gets the gradient of the same problem run alone,
  bit for bit;
- the reverse sweep that takes the last segment from the recorded forward's
  tape equals one that replays it.
"""
import numpy as np
import pytest

from adjoint_host.harness import Host, build
from cuteflow_amd.adjoint import AdjointRun, GaugeLevels, zone_map
from cuteflow_amd.solver import BatchSolver, Solver

from cases import dam_case, river_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hs(tmp_path_factory):
    return build(tmp_path_factory.mktemp("hs_gpu"))


def _same(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def _gauges(m, k, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(m.nelt, size=k, replace=False)).astype(np.int32)


CASES = {
    "river_waf": lambda: river_case(30, 8, seed=21),
    "river_hllc_f2": lambda: river_case(30, 8, seed=22, iflux=1, fricimplic=2),
    "dam_hllc": lambda: dam_case(14, seed=5),
    "river_riadh_f1": lambda: river_case(30, 8, seed=23, iflux=2, fricimplic=1),
    "river_muscl_l3": lambda: river_case(30, 8, seed=24, iflux=4, ilimiteur=3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_tlm_and_adjoint_match_host(hs, name):
    m, p, U0 = CASES[name]()
    nsteps = 60
    print(f"case {name} nelt {m.nelt}", flush=True)
    H = Host(hs, m, p)
    Uh, hist_h, done = H.forward(U0, nsteps)
    assert done == nsteps
    rng = np.random.default_rng(11)
    dn = rng.normal(size=m.nelt) * 0.01
    dU0 = rng.normal(size=(m.nelt, 3)) * 0.01
    dU0[U0[:, 0] <= p.tolisec] = 0.0
    lam = rng.normal(size=(m.nelt, 3))
    dUh, _ = H.tlm(dn, dU0)
    gnh, gUh, bad = H.adjoint(lam)
    H.close()
    assert bad == 0
    with Solver(m, p) as s:
        run = AdjointRun(s, np.zeros(0, np.int32), sample_every=7, checkpoint_every=16)
        run.forward(nsteps, U0=[U0])
        assert _same(s.get_state(), Uh)
        t = run.tlm([dn], [dU0])[0]
        assert _same(t["state"], Uh)
        assert _same(t["d_state"], dUh)
        gn, gU = run.backward(None, None, [lam])
    scale_n = np.max(np.abs(gnh)) + 1e-300
    scale_u = np.max(np.abs(gUh)) + 1e-300
    assert np.max(np.abs(gn[0] - gnh)) <= 1e-12 * scale_n
    assert np.max(np.abs(gU[0] - gUh)) <= 1e-12 * scale_u


def test_gpu_dot_product_all_seeds():
    m, p, U0 = river_case(40, 10, seed=31)
    nsteps = 90
    g = _gauges(m, 12, 1)
    rng = np.random.default_rng(5)
    with Solver(m, p) as s:
        run = AdjointRun(s, g, sample_every=10, checkpoint_every=25, want_hmax=True)
        etas, hmaxs = run.forward(nsteps, U0=[U0])
        nsamp = run.nsamp
        assert nsamp == 9 and etas[0].shape == (9, 12)
        dn = rng.normal(size=m.nelt) * 0.005
        dU0 = rng.normal(size=(m.nelt, 3)) * 0.01
        dU0[U0[:, 0] <= p.tolisec] = 0.0
        t = run.tlm([dn], [dU0])[0]
        # the TLM values at the recorded n are the recorded samples
        assert _same(t["eta"], etas[0])
        ge = rng.normal(size=(nsamp, 12))
        gh = rng.normal(size=m.nelt) * (hmaxs[0] > p.tolisec)
        gs = rng.normal(size=(m.nelt, 3))
        gn, gU = run.backward([ge], [gh], [gs])
    lhs = np.sum(ge * t["d_eta"]) + np.sum(gh * t["d_hmax"]) + np.sum(gs * t["d_state"])
    rhs = np.sum(gn[0] * dn) + np.sum(gU[0] * dU0)
    scale = np.sum(np.abs(ge * t["d_eta"])) + np.sum(np.abs(gh * t["d_hmax"])) + np.sum(np.abs(gs * t["d_state"]))
    assert abs(lhs - rhs) <= 1e-11 * scale, (lhs, rhs, scale)


@pytest.mark.parametrize("iflux", [1, 2, 3, 4])
def test_gpu_zonal_gradient_matches_tlm_and_central_differences(iflux):
    """Gauge misfit over 80 steps, gradient with respect to 4 zonal n through
    the autograd Function, against (a) the tangent-linear directional
    derivative along each zone (exact derivative of the same map: 1e-12) and
    (b) central differences of the forward with the recorded dt at relative
    step 1e-6.  The WAF and MUSCL limiters are piecewise (flux_riemann.cuf:
    430-524): a step that straddles a kink of one face's limiter at one step
    changes the measured slope, so for them the best of the steps 1e-6,
    1e-7, 1e-8 is held to the bar; the first-order solvers (iflux 1, 2) meet
    it at 1e-6.  The tolerance is 1e-6 of
    the zone's derivative plus 1e-9 of the largest (round-off floor of a
    difference quotient of a loss of order 1)."""
    import torch

    m, p, U0 = river_case(40, 10, seed=41, iflux=iflux)
    nz, nsteps = 4, 80
    xc = m.centroids()[0]
    zone = np.minimum((xc / xc.max() * nz).astype(np.int64), nz - 1)
    g = _gauges(m, 10, 2)
    dev = torch.device("cuda")
    with Solver(m, p) as s:
        run = AdjointRun(s, g, sample_every=10, checkpoint_every=32)
        Z = zone_map(zone, nz, device=dev)
        Zn = Z.cpu().numpy()
        nz0 = np.array([0.03, 0.045, 0.06, 0.04])
        obs = run.forward(nsteps, U0=[U0], manning=[Zn @ (nz0 * 0.8)])[0][0].reshape(-1)
        n_z = torch.tensor(nz0, dtype=torch.float64, device=dev, requires_grad=True)
        eta = GaugeLevels.apply(Z @ n_z, run, nsteps, [U0])
        loss = ((eta - torch.as_tensor(obs, device=dev)) ** 2).sum()
        loss.backward()
        grad = n_z.grad.detach().cpu().numpy()

        def loss_at(nzp):
            run.set_manning([Zn @ nzp])
            e = run.tlm()[0]["eta"].reshape(-1)
            return float(np.sum((e - obs) ** 2))

        tlm = np.zeros(nz)
        run.set_manning([Zn @ nz0])
        for z in range(nz):
            t = run.tlm([Zn[:, z].copy()])[0]
            tlm[z] = np.sum(2.0 * (t["eta"].reshape(-1) - obs) * t["d_eta"].reshape(-1))
        steps = (1e-6,) if iflux <= 2 else (1e-6, 1e-7, 1e-8)
        fd = np.zeros((len(steps), nz))
        for k, eps in enumerate(steps):
            for z in range(nz):
                up, dn = nz0.copy(), nz0.copy()
                up[z] *= 1.0 + eps
                dn[z] *= 1.0 - eps
                fd[k, z] = (loss_at(up) - loss_at(dn)) / (2.0 * eps * nz0[z])
    print("adjoint", grad, "tlm", tlm, "fd", fd, flush=True)
    assert np.all(np.abs(grad - tlm) <= 1e-12 * np.abs(tlm) + 1e-15 * np.max(np.abs(tlm)))
    tol = 1e-6 * np.abs(grad) + 1e-9 * np.max(np.abs(grad))
    best = np.min(np.abs(fd - grad[None, :]), axis=0)
    assert np.all(best <= tol), (grad, fd)


def test_gpu_batch_gradients_equal_single_runs():
    cases = [river_case(30, 8, seed=51), river_case(20, 6, seed=52, q_peak=300.0), river_case(36, 8, seed=53)]
    gauges = [_gauges(c[0], 6, 60 + k) for k, c in enumerate(cases)]
    nsteps = 50
    rng = np.random.default_rng(9)
    seeds = [rng.normal(size=(5, 6)) for _ in cases]
    single = []
    for (m, p, U0), g, ge in zip(cases, gauges, seeds):
        with Solver(m, p) as s:
            run = AdjointRun(s, g, sample_every=10, checkpoint_every=20)
            eta, _ = run.forward(nsteps, U0=[U0])
            gn, gU = run.backward([ge])
            single.append((eta[0], gn[0], gU[0]))
    with BatchSolver([c[0] for c in cases], [c[1] for c in cases]) as bs:
        run = AdjointRun(bs, gauges, sample_every=10, checkpoint_every=20)
        etas, _ = run.forward(nsteps, U0=[c[2] for c in cases])
        gns, gUs = run.backward(seeds)
    for b, (eta, gn, gU) in enumerate(single):
        assert _same(etas[b], eta), b
        assert _same(gns[b], gn), b
        assert _same(gUs[b], gU), b


def test_gpu_lbfgs_calibration_reduces_misfit():
    """torch.optim.LBFGS on 4 zonal n from a +20 % start: the gauge misfit
    against synthetic observations of the true n drops by orders of
    magnitude within a few iterations."""
    import torch

    m, p, U0 = river_case(40, 10, seed=71, iflux=1)
    nz, nsteps = 4, 60
    xc = m.centroids()[0]
    zone = np.minimum((xc / xc.max() * nz).astype(np.int64), nz - 1)
    dev = torch.device("cuda")
    with Solver(m, p) as s:
        run = AdjointRun(s, _gauges(m, 16, 3), sample_every=10, checkpoint_every=30)
        Z = zone_map(zone, nz, device=dev)
        truth = np.array([0.03, 0.04, 0.05, 0.035])
        obs = torch.as_tensor(run.forward(nsteps, U0=[U0], manning=[Z.cpu().numpy() @ truth])[0][0].reshape(-1),
                              device=dev)
        n_z = torch.tensor(truth * 1.2, dtype=torch.float64, device=dev, requires_grad=True)
        with torch.no_grad():
            l0 = ((GaugeLevels.apply(Z @ n_z, run, nsteps, [U0]) - obs) ** 2).sum().item()
        opt = torch.optim.LBFGS([n_z], lr=1.0, max_iter=12, history_size=8, tolerance_grad=1e-12,
                                tolerance_change=1e-14, line_search_fn="strong_wolfe")
        losses = []

        def closure():
            opt.zero_grad()
            eta = GaugeLevels.apply(Z @ n_z, run, nsteps, [U0])
            loss = ((eta - obs) ** 2).sum() / l0  # misfit relative to the start
            loss.backward()
            losses.append(loss.item())
            return loss

        opt.step(closure)
    print("losses", losses, "n", n_z.detach().cpu().numpy(), flush=True)
    assert min(losses) < 1e-3


def test_gpu_reverse_tile_skipping_is_exact():
    """A reach with a dry floodplain: tiles the recorded forward skipped are
    skipped by the replay, reverse and gather kernels too.  Gradients with
    256- and 64-cell tiles (and 128-cell tiles in XCD-contiguous order)
    equal those of a run that processes every tile (CF_OPT_NO_TILE_SKIP),
    value for value.  The maximum depths (whose sampling passes over the
    tiles that did not run since the previous sample) equal the maxima of
    the states a plain run reaches at the sample steps, bit for bit."""
    from cuteflow_amd import _lib
    from cuteflow_amd import scenarios as SC

    sc = SC.reach(40_000, seed=81, iflux=1, q_peak=800.0, t_ramp=600.0)
    m, p, U0 = sc.mesh, sc.params, sc.state
    nsteps, every = 120, 10
    g = _gauges(m, 20, 4)
    rng = np.random.default_rng(12)
    ge = rng.normal(size=(12, 20))
    gh = rng.normal(size=m.nelt)
    gs = rng.normal(size=(m.nelt, 3))
    out = []
    for opt in (0, _lib.CF_OPT_TILE_64, _lib.CF_OPT_TILE_128 | _lib.CF_OPT_XCD_CONTIG, _lib.CF_OPT_NO_TILE_SKIP):
        with Solver(m, p, options=opt) as s:
            s.set_state(U0)
            hmax_ref = None
            for _ in range(nsteps // every):
                s.step(every)
                h = s.get_state()[:, 0]
                hmax_ref = h.copy() if hmax_ref is None else np.where(h > hmax_ref, h, hmax_ref)
            frac = s.last_run_stats()["active_fraction"]
            run = AdjointRun(s, g, sample_every=every, checkpoint_every=50, want_hmax=True)
            eta, hm = run.forward(nsteps, U0=[U0])
            assert _same(hm[0], hmax_ref), opt
            gn, gU = run.backward([ge], [gh], [gs])
            out.append((frac, eta[0], gn[0], gU[0], hm[0]))
    print("active fractions", [o[0] for o in out], flush=True)
    assert all(o[0] < 1.0 for o in out[:3]) and out[3][0] == 1.0  # some tiles were skipped
    for k in (0, 1, 2):
        assert _same(out[k][1], out[3][1])
        assert _same(out[k][4], out[3][4])
        assert np.array_equal(out[k][2], out[3][2])
        assert np.array_equal(out[k][3], out[3][3])


def test_gpu_hmax_sampling_long_interval():
    """Samples 70 steps apart (the activity check of a tile reads 70 rows,
    more than one wave's lanes): the maximum depths equal the maxima of the
    states a plain run reaches at the sample steps, bit for bit."""
    from cuteflow_amd import scenarios as SC

    sc = SC.reach(40_000, seed=82, iflux=1, q_peak=800.0, t_ramp=600.0)
    m, p, U0 = sc.mesh, sc.params, sc.state
    nsteps, every = 210, 70
    with Solver(m, p) as s:
        s.set_state(U0)
        hmax_ref = None
        for _ in range(nsteps // every):
            s.step(every)
            h = s.get_state()[:, 0]
            hmax_ref = h.copy() if hmax_ref is None else np.where(h > hmax_ref, h, hmax_ref)
        assert s.last_run_stats()["active_fraction"] < 1.0
        run = AdjointRun(s, _gauges(m, 5, 2), sample_every=every, checkpoint_every=100, want_hmax=True)
        _, hm = run.forward(nsteps, U0=[U0])
    assert _same(hm[0], hmax_ref)


def test_gpu_reverse_from_forward_tape_equals_replay():
    """The reverse sweep after a recorded forward reads the last segment's
    states and residuals as the forward wrote them; a second sweep over the
    same record replays every segment, and a record in one segment is all
    tape.  The three give the same gradients, value for value."""
    from cuteflow_amd import scenarios as SC

    sc = SC.reach(40_000, seed=81, iflux=1, q_peak=800.0, t_ramp=600.0)
    m, p, U0 = sc.mesh, sc.params, sc.state
    nsteps = 110
    g = _gauges(m, 20, 4)
    rng = np.random.default_rng(13)
    ge = rng.normal(size=(11, 20))
    gh = rng.normal(size=m.nelt)
    gs = rng.normal(size=(m.nelt, 3))
    out = []
    for K, sweeps in ((50, 2), (200, 1)):  # segments of 50, 50 and 10 steps; one segment
        with Solver(m, p) as s:
            run = AdjointRun(s, g, sample_every=10, checkpoint_every=K, want_hmax=True)
            run.forward(nsteps, U0=[U0])
            for _ in range(sweeps):
                gn, gU = run.backward([ge], [gh], [gs])
                out.append((gn[0], gU[0]))
    assert np.any(out[0][0] != 0.0)
    for k in (1, 2):
        assert np.array_equal(out[k][0], out[0][0])
        assert np.array_equal(out[k][1], out[0][1])


def test_gpu_branch_masks():
    """Diagnostic masks of tangent-linear runs: recording them changes no
    value or tangent (the recording kernel's arithmetic is the plain one's);
    a run at the recorded n takes every branch as recorded (no flips of
    either kind); a run at perturbed n on a reach that wets and dries takes
    some differently, and every changed wet/dry decision is also a changed
    branch.  A new recorded forward turns the masks off."""
    m, p, U0 = river_case(30, 8, seed=25, iflux=2, fricimplic=1)
    nsteps = 80
    rng = np.random.default_rng(3)
    dn = rng.normal(size=m.nelt) * 0.01
    with Solver(m, p) as s:
        run = AdjointRun(s, _gauges(m, 6, 4), sample_every=8, checkpoint_every=20)
        run.forward(nsteps, U0=[U0])
        plain = run.tlm([dn])[0]
        run.set_mask_mode(3)
        rec = run.tlm([dn])[0]
        for k in ("eta", "d_eta", "state", "d_state"):
            assert _same(plain[k], rec[k]), k
        run.set_mask_mode(4)
        run.tlm([dn])
        f0, f1 = run.flips(0)[0], run.flips(1)[0]
        assert f0.sum() == 0 and f1.sum() == 0
        run.set_manning([m.manning * 1.5])
        run.tlm([dn])
        f0, f1 = run.flips(0)[0], run.flips(1)[0]
        assert f1.sum() > 0 and f0.sum() <= f1.sum()
        assert np.all(f1[f0 > 0] > 0)
        run.set_manning([m.manning])
        run.forward(nsteps, U0=[U0])
        with pytest.raises(Exception):
            run.set_mask_mode(4)            # masks of the former record are gone
        run.set_mask_mode(0)
