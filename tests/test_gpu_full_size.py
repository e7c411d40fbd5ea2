"""The configurations at their full size on the GPU.

- River reach, 2,048,000 cells (3200 x 320 quads), WAF + superbee +
  semi-implicit friction, stored in RCM and in random cell order: 100 steps
  from the initial state, bit-identical to the oracle (states and the dt
  sequence); the two orders also give the same bits as each other.
- Zonal calibration, 100k cells, 8 zones, 20 gauges, levels every 10 steps
  over 2000 steps, checkpoints every 100 steps, first-order schemes (HLLC and
  the reference's calibration scheme, iflux 2): the adjoint zonal gradient
  against tangent-linear directional derivatives (1e-12), against central
  differences at the recorded dt with step 1e-7 (1e-6 on every zone whose
  perturbed runs take no branch differently; a zone past 1e-6 has such
  branch flips), and 20 L-BFGS iterations recover every zone's n.
- Per-cell Manning fields, 16 members on a 1M-cell reach, 3000 steps,
  checkpoints every 150 steps, gauges + flood-extent loss: the adjoint
  against tangent-linear directional derivatives along 4 random directions
  (1e-12).
"""
import numpy as np
import pytest

from cuteflow_amd import calibration as CAL
from cuteflow_amd import scenarios as SC
from cuteflow_amd.adjoint import AdjointRun
from cuteflow_amd.solver import BatchSolver, Solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def reach_runs():
    """(order -> (final state in generation order, dt history)) of the GPU runs."""
    return {}


@pytest.mark.parametrize("order,options", [("rcm", 24), ("random", 0)])
def test_river_reach_full_size_matches_oracle(order, options, reach_runs):
    nsteps = 100
    m, p, U0, perm = SC.river_reach(3200, 320, order)
    print(f"river reach: nelt={m.nelt} order={order} options={options} steps={nsteps}", flush=True)
    with Solver(m, p, options=options) as s:
        s.set_state(U0)
        hist, done, _ = s.step(nsteps)
        U = s.get_state()
    assert done == nsteps
    ref = O.run(O.Prepared(m), p, U0, nsteps)
    assert ref["steps"] == nsteps
    assert np.array_equal(hist, ref["dt_hist"]), "dt sequence differs"
    assert np.array_equal(U, ref["U"]), f"state differs: max {np.max(np.abs(U - ref['U']))}"
    wet = int(np.sum(U[:, 0] > p.tolisec))
    assert wet > 100_000                                    # the channel is wet, the floodplain dry
    nat = np.empty_like(U)
    nat[perm] = U
    reach_runs[order] = (nat, hist)
    if len(reach_runs) == 2:
        (a, ha), (b, hb) = reach_runs.values()
        assert np.array_equal(a, b) and np.array_equal(ha, hb)


@pytest.mark.parametrize("iflux", [1, 2])
def test_calibration_full_size_gradient(iflux):
    nsteps, every, K = 2000, 10, 100
    sc = SC.calibration(100_000, ngauges=20, iflux=iflux)
    N = sc.mesh.nelt
    Z = np.zeros((N, 8))
    Z[np.arange(N), sc.zone] = 1.0
    print(f"calibration: nelt={N} iflux={iflux} steps={nsteps}", flush=True)
    with Solver(sc.mesh, sc.params, options=24) as s:
        run = AdjointRun(s, sc.gauges, sample_every=every, checkpoint_every=K)
        obs = run.forward(nsteps, U0=[sc.state], manning=[Z @ sc.n_zones])[0][0]
        n0 = sc.n_zones * 1.2
        eta = run.forward(nsteps, U0=[sc.state], manning=[Z @ n0])[0][0]
        gn, _ = run.backward([2.0 * (eta - obs)])
        g = Z.T @ gn[0]
        g_tlm = CAL.zonal_tlm_gradient(run, Z, obs)
        assert np.all(np.abs(g - g_tlm) <= 1e-12 * np.abs(g_tlm)), (g, g_tlm)
        # every zone is observable: gradients within three orders of magnitude
        assert np.min(np.abs(g)) > 1e-3 * np.max(np.abs(g))
        g_fd, clamp_flips, branch_flips = CAL.zonal_fd_gradient(run, Z, n0, obs, eps=1e-7, count_flips=True)
        rel = np.abs(g - g_fd) / np.abs(g_fd)
        print(f"calibration: fd rel {rel.tolist()} clamp flips {clamp_flips.tolist()} "
              f"branch flips {branch_flips.tolist()}", flush=True)
        assert np.all(clamp_flips <= branch_flips)
        assert np.all((rel <= 1e-6) | (branch_flips > 0)), (rel, branch_flips)
        nz, hist, _ = CAL.zonal_lbfgs(run, sc.zone, n0, obs, nsteps, sc.state, iters=20)
        err = np.abs(nz - sc.n_zones) / sc.n_zones
        print(f"calibration: L-BFGS relative errors {err.tolist()}", flush=True)
        assert np.all(err < 1e-4), err


def test_field_ensemble_full_size_tlm():
    nsteps, every, K = 3000, 10, 150
    mem = SC.field_ensemble(16, 1_000_000, ngauges=50, iflux=1)
    B = len(mem)
    print(f"field ensemble: {B} x {mem[0].mesh.nelt} cells, steps={nsteps}", flush=True)
    with BatchSolver([e.mesh for e in mem], [e.params for e in mem], options=24) as bs:
        U0 = [e.state for e in mem]
        surf = [bs.geometry(b)["surf"] for b in range(B)]
        run = AdjointRun(bs, [e.gauges for e in mem], sample_every=every, checkpoint_every=K, want_hmax=True)
        obs, hobs = run.forward(nsteps, U0=U0, manning=[e.mesh.manning for e in mem])
        a_obs = [CAL.soft_extent(hobs[b], surf[b])[0] for b in range(B)]
        scale = [max(1.0, 0.01 * a) for a in a_obs]
        n0 = [1.2 * e.mesh.manning for e in mem]
        etas, hm = run.forward(nsteps, U0=U0, manning=n0)
        ge, gh = [], []
        for b in range(B):
            ge.append(2.0 * (etas[b] - obs[b]))
            gh.append(CAL.extent_loss(hm[b], surf[b], a_obs[b], scale[b])[1])
        gn, _ = run.backward(ge, gh)
        rng = np.random.default_rng(3)
        for d in range(4):
            dn = [n0[b] * rng.normal(0.0, 0.1, n0[b].shape) for b in range(B)]
            tl = run.tlm(dn)
            dl_tlm = sum(float(np.sum(ge[b] * tl[b]["d_eta"])) + float(np.sum(gh[b] * tl[b]["d_hmax"]))
                         for b in range(B))
            dl_adj = float(sum(np.sum(gn[b] * dn[b]) for b in range(B)))
            print(f"field ensemble: direction {d}: tlm {dl_tlm:.17g} adjoint {dl_adj:.17g}", flush=True)
            assert abs(dl_adj - dl_tlm) <= 1e-12 * abs(dl_tlm)
