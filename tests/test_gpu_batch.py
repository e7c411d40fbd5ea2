"""Batched instances (cf_batch_*): each instance of a batch must give the
same bits as the oracle run of that instance alone, whatever the other
instances do (different sizes, boundary conditions, end times, failures,
state resets), and runs replayed from a captured graph must match runs
launched one kernel at a time."""
import os

import numpy as np
import pytest

from cuteflow_amd.solver import BatchSolver, Solver
from oracle import oracle as O

from cases import dam_case, river_case

pytestmark = pytest.mark.gpu


def _same(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def _hllc_batch():
    """HLLC without friction; walls/transmissive dam breaks of three sizes,
    one with an early end time, one all dry (fails after one step), and a
    river reach with an inflow hydrograph and a fixed outlet level."""
    cases = []
    m, p, U = dam_case(16)
    cases.append((m, p, U))
    m, p, U = dam_case(10, seed=1)
    p.ts = 0.5
    cases.append((m, p, U))
    m, p, U = dam_case(8, seed=2)
    U[:] = 0.0
    cases.append((m, p, U))
    m, p, U = river_case(40, 8, iflux=1, friction=0)
    cases.append((m, p, U))
    return cases


def test_batch_mixed_instances_match_oracle():
    cases = _hllc_batch()
    nsteps = 120
    refs = [O.run(O.Prepared(m), p, U, nsteps) for m, p, U in cases]
    with BatchSolver([c[0] for c in cases], [c[1] for c in cases]) as bs:
        for b, (_, _, U) in enumerate(cases):
            bs.set_state(b, U)
        hist, done, status = bs.step(nsteps)
        for b, ref in enumerate(refs):
            print(f"instance {b}: nelt {cases[b][0].nelt} done {done[b]} status {status[b]}", flush=True)
            assert done[b] == ref["steps"], (b, done[b], ref["steps"])
            assert (status[b] == 0) == (ref["status"] == 0), (b, status[b], ref["status"])
            assert _same(hist[b], ref["dt_hist"]), b
            assert _same(bs.get_state(b), ref["U"]), b
            assert bs.time(b)[0] == ref["tc"], b
    assert refs[1]["steps"] < nsteps  # stopped at its own ts
    assert refs[2]["steps"] == 1 and refs[2]["status"] != 0  # all dry: dt error


def test_batch_waf_friction_uneven_sizes_and_manning():
    cases = [river_case(40, 8, seed=3), river_case(24, 6, seed=4, q_peak=300.0),
             river_case(60, 12, seed=5, ilimiteur=2)]
    # per-cell Manning on the second instance
    m1 = cases[1][0]
    rng = np.random.default_rng(7)
    n1 = 0.02 + 0.04 * rng.random(m1.nelt)
    m1.manning = n1.copy()
    nsteps = 150
    refs = [O.run(O.Prepared(m), p, U, nsteps) for m, p, U in cases]
    with BatchSolver([c[0] for c in cases], [c[1] for c in cases]) as bs:
        for b, (_, _, U) in enumerate(cases):
            bs.set_state(b, U)
        hist, done, status = bs.step(nsteps)
        for b, ref in enumerate(refs):
            print(f"instance {b}: nelt {cases[b][0].nelt} done {done[b]}", flush=True)
            assert done[b] == ref["steps"] == nsteps
            assert _same(hist[b], ref["dt_hist"]), b
            assert _same(bs.get_state(b), ref["U"]), b


def test_batch_reset_one_instance_mid_run():
    """Resetting instance 1 restarts its clock; instance 0 goes on."""
    cases = [river_case(30, 6, seed=8), river_case(20, 6, seed=9)]
    with BatchSolver([c[0] for c in cases], [c[1] for c in cases]) as bs:
        for b, (_, _, U) in enumerate(cases):
            bs.set_state(b, U)
        bs.step(30)
        bs.set_state(1, cases[1][2])
        assert bs.time(1)[0] == 0.0
        bs.step(25)
        bs.step(15)
        r0 = O.run(O.Prepared(cases[0][0]), cases[0][1], cases[0][2], 70)
        r1 = O.run(O.Prepared(cases[1][0]), cases[1][1], cases[1][2], 40)
        assert _same(bs.get_state(0), r0["U"])
        assert _same(bs.get_state(1), r1["U"])
        assert bs.time(0)[0] == r0["tc"] and bs.time(1)[0] == r1["tc"]


def test_graph_replay_matches_direct_launches():
    """Runs of a repeated length replay one captured graph; the same runs
    launched kernel by kernel (CUTEFLOW_NO_GRAPHS=1) give the same bits, and
    both equal one oracle run of the total length."""
    m, p, U0 = river_case(40, 8, seed=11)
    out = []
    for env in ("0", "1"):
        os.environ["CUTEFLOW_NO_GRAPHS"] = env
        try:
            with Solver(m, p) as s:
                s.set_state(U0)
                hists = [s.step(20)[0][:-1] for _ in range(5)]
                out.append((np.concatenate(hists), s.get_state(), s.tc))
        finally:
            os.environ.pop("CUTEFLOW_NO_GRAPHS", None)
    ref = O.run(O.Prepared(m), p, U0, 100)
    for hist, U, tc in out:
        assert _same(hist, ref["dt_hist"][:-1])
        assert _same(U, ref["U"])
        assert tc == ref["tc"]


def test_batch_rejects_mixed_schemes():
    m0, p0, _ = dam_case(8)
    m1, p1, _ = river_case(20, 6)
    with pytest.raises(RuntimeError):
        BatchSolver([m0, m1], [p0, p1])


def test_autotune_picks_a_candidate_with_identical_results():
    """cf_batch_autotune times every schedule candidate and returns one of
    them; the batch built with the tuned options gives the same bits as the
    default schedule (and the oracle), on instances of uneven sizes."""
    from cuteflow_amd import _lib
    from cuteflow_amd.solver import autotune, tune_candidate_options
    cases = [river_case(60, 12, seed=21, q_peak=1500.0, t_ramp=30.0), river_case(30, 8, seed=22),
             river_case(90, 10, seed=23, q_peak=800.0)]
    meshes, params, states = [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases]
    best, table = autotune(meshes, params, states, nsteps=40)
    print("autotune:", best, table, flush=True)
    assert len(table) == _lib.CF_TUNE_NCAND
    assert [o for o, _ in table] == [tune_candidate_options(k) for k in range(_lib.CF_TUNE_NCAND)]
    assert all(t is not None and t > 0.0 for _, t in table)
    k_best = min(range(len(table)), key=lambda k: table[k][1])
    if k_best & 2:  # skipping off: a skipping candidate within 3 % is preferred
        near = [k for k in range(len(table)) if not k & 2 and table[k][1] <= 1.03 * table[k_best][1]]
        if near:
            k_best = min(near, key=lambda k: table[k][1])
    assert best == table[k_best][0]
    nsteps = 200
    refs = [O.run(O.Prepared(m), p, U, nsteps) for m, p, U in cases]
    for options in (best, 0, _lib.CF_OPT_XCD_CONTIG | _lib.CF_OPT_TILE_64):
        with BatchSolver(meshes, params, options=options) as bs:
            for b, U in enumerate(states):
                bs.set_state(b, U)
            hist, done, status = bs.step(nsteps)
            for b, ref in enumerate(refs):
                assert done[b] == ref["steps"], (options, b)
                assert _same(hist[b], ref["dt_hist"]), (options, b)
                assert _same(bs.get_state(b), ref["U"]), (options, b)


def test_autotune_rejects_bad_arguments():
    from cuteflow_amd.solver import autotune
    m, p, U = river_case(20, 6)
    with pytest.raises(ValueError):
        autotune([m], [p], [U[:-1]], nsteps=5)
    with pytest.raises(RuntimeError):
        autotune([m], [p], [U], nsteps=0)


@pytest.mark.parametrize("iflux", [1, 3])
def test_batch_with_tiny_instances(iflux):
    """Instances of 2 and 66 triangles (smaller than a tile, one tile and two
    cells) batched with a river reach, 256- and 64-cell tiles: each instance
    gives the oracle's bits for that instance alone."""
    from cuteflow_amd import _lib
    from test_gpu_parity import _tiny_case
    cases = [_tiny_case(1, 1, iflux, seed=3), _tiny_case(33, 1, iflux, seed=4),
             river_case(30, 8, seed=25, iflux=iflux), _tiny_case(2, 2, iflux, seed=5)]
    nsteps = 40
    refs = [O.run(O.Prepared(m), p, U, nsteps) for m, p, U in cases]
    for options in (0, _lib.CF_OPT_TILE_64 | _lib.CF_OPT_XCD_CONTIG):
        with BatchSolver([c[0] for c in cases], [c[1] for c in cases], options=options) as bs:
            for b, (_, _, U) in enumerate(cases):
                bs.set_state(b, U)
            hist, done, status = bs.step(nsteps)
            for b, ref in enumerate(refs):
                assert done[b] == ref["steps"], (options, b)
                assert _same(hist[b], ref["dt_hist"]), (options, b)
                assert _same(bs.get_state(b), ref["U"]), (options, b)
