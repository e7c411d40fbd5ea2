"""Parity of the HIP time step (through the C ABI) with the CPU oracle.

The bar: bit-identical states, dt sequences and wet/dry masks on the same
seeded inputs (the kernels and the oracle evaluate the same FP64 operations
in the same order, without FMA contraction).  The parity tolerances
(1e-10 relative L-inf on h, hu, hv; 1e-12 on dt) are asserted as well, so a
failure message shows which bar broke.
"""
import numpy as np
import pytest

from cuteflow_amd import mesh as M
from cuteflow_amd.params import Params
from cuteflow_amd.solver import Solver
from oracle import oracle as O

from cases import dam_case, lake_case, river_case

pytestmark = pytest.mark.gpu

RTOL_STATE = 1e-10
RTOL_DT = 1e-12


def _compare(U_gpu, U_cpu, dt_gpu, dt_cpu, tolisec):
    assert dt_gpu.shape == dt_cpu.shape, (dt_gpu.shape, dt_cpu.shape)
    rel_dt = np.max(np.abs(dt_gpu - dt_cpu) / np.abs(dt_cpu))
    assert rel_dt <= RTOL_DT, f"dt sequence differs: {rel_dt}"
    for k in range(3):
        scale = max(np.max(np.abs(U_cpu[:, k])), 1e-300)
        err = np.max(np.abs(U_gpu[:, k] - U_cpu[:, k])) / scale
        assert err <= RTOL_STATE, f"component {k}: rel Linf {err}"
    wet_g = U_gpu[:, 0] > tolisec
    wet_c = U_cpu[:, 0] > tolisec
    assert np.array_equal(wet_g, wet_c), f"{np.sum(wet_g != wet_c)} wet/dry mismatches"
    # stronger than the tolerances: the same bits
    assert np.array_equal(dt_gpu.view(np.int64), dt_cpu.view(np.int64))
    assert np.array_equal(U_gpu.view(np.int64), U_cpu.view(np.int64)), \
        f"{np.sum(U_gpu != U_cpu)} values differ in the last bits"


def _run_both(m, p, U0, nsteps, perm=None):
    prep = O.Prepared(m)
    ref = O.run(prep, p, U0, nsteps)
    if perm is not None:
        inv = np.empty_like(perm)
        inv[perm] = np.arange(len(perm))
        mg, U0g = m.permuted(perm), U0[perm]
    else:
        mg, U0g = m, U0
    with Solver(mg, p) as s:
        s.set_state(U0g)
        dt0 = s.dt
        hist, done, st = s.step(nsteps, raise_on_error=False)
        U = s.get_state()
        tc = s.tc
    if perm is not None:
        U = U[inv]
    return ref, U, np.asarray(hist), done, st, dt0, tc


def test_geometry_bitwise():
    m, p, _ = river_case()
    prep = O.Prepared(m)
    with Solver(m, p) as s:
        g = s.geometry()
    assert np.array_equal(g["surf"], prep.surf)
    assert np.array_equal(g["zm"], prep.zm)
    assert np.array_equal(g["cotemin_arr"], prep.cot)
    assert g["cotemin"] == prep.cotemin


def test_pow_bitwise():
    from cuteflow_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(0)
    x = np.concatenate([10.0 ** rng.uniform(-6, 3, 200000), rng.uniform(1e-3, 20, 100000)])
    for y in (-1.0 / 3.0, -4.0 / 3.0, 1.0 / 3.0, 0.5, 2.5):
        out = np.zeros_like(x)
        _lib.check(lib.cf_debug_pow(len(x), _lib.dptr(x), y, _lib.dptr(out)), "cf_debug_pow")
        ref = np.array([O.pow_(v, y) for v in x[:20000]])
        assert np.array_equal(out[:20000], ref)
        ulp = np.abs(out - np.power(x, y)) / np.spacing(np.power(x, y))
        assert ulp.max() <= 2.0


def test_initial_dt():
    m, p, U0 = dam_case(30)
    prep = O.Prepared(m)
    with Solver(m, p) as s:
        s.set_state(U0)
        assert s.dt == O.cfl_dt(prep, p, U0)


@pytest.mark.parametrize("n,steps", [(20, 200), (100, 1000)])
def test_dam_break_hllc(n, steps):
    """Circular dam break, HLLC, walls; (100, 1000) is the full 20k-cell case."""
    m, p, U0 = dam_case(n)
    ref, U, hist, done, st, _, tc = _run_both(m, p, U0, steps)
    assert st == 0 and ref["status"] == 0 and done == steps == ref["steps"]
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)
    assert tc == ref["tc"]


@pytest.mark.parametrize("fricimplic", [1, 2])
def test_river_waf_friction(fricimplic):
    m, p, U0 = river_case(fricimplic=fricimplic)
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 300)
    assert st == 0 and ref["status"] == 0 and done == 300
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)
    assert np.sum(U[:, 0] > p.tolisec) > 0


def test_river_explicit_friction_blowup_matches():
    """Explicit friction (fricimplic 0) on the shallow wetting front is
    unstable: dt collapses below 1e-12 after a few steps.  Both sides must
    stop at the same step with the same state and dt sequence."""
    m, p, U0 = river_case(fricimplic=0)
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 300)
    assert ref["status"] == -1 and st == -3
    assert done == ref["steps"] < 300
    assert np.array_equal(hist, ref["dt_hist"])
    assert np.array_equal(U, ref["U"])


def test_dam_break_explicit_friction():
    m, p, U0 = dam_case(24)
    p.friction, p.fricimplic = 1, 0
    m.manning[:] = 0.03
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 300)
    assert st == 0 and ref["status"] == 0 and done == 300
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


@pytest.mark.parametrize("ilim", [0, 1, 2, 3, 4])
def test_waf_limiters(ilim):
    m, p, U0 = river_case(ilimiteur=ilim, nx=30, ny=9)
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 150)
    assert st == 0 and done == 150
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


@pytest.mark.parametrize("fricimplic", [0, 1, 2])
def test_riadh_river(fricimplic):
    """iflux 2 (solver_Riemann_riadh) on the wetting reach, every friction
    mode (explicit friction on the dam break: it is unstable on the reach's
    shallow front)."""
    if fricimplic == 0:
        m, p, U0 = dam_case(24)
        p.iflux, p.friction, p.fricimplic = 2, 1, 0
        m.manning[:] = 0.03
    else:
        m, p, U0 = river_case(iflux=2, fricimplic=fricimplic)
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 300)
    assert st == 0 and ref["status"] == 0 and done == 300
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


def test_riadh_reference_canal_runs():
    """The divergent canal's two reference GPU runs (iflux 2): the GPU gives
    the oracle's bits over the whole logged length (16000 and 15000 steps),
    and the oracle reproduces the reference's logged outflow discharges."""
    import json
    import os

    from cuteflow_amd import mesh as Mm
    from cuteflow_amd.params import read_donnees
    from cuteflow_amd.solver import initial_state

    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    m = Mm.read_text_mesh(os.path.join(golden, "divergent_canal.mesh"))
    exp = json.load(open(os.path.join(golden, "divergent_canal.expected.json")))
    prep = O.Prepared(m)
    for log_name, n in (("outfile_manning_min.out", 0.01), ("outfile_manning_max.out", 0.05)):
        log = exp["logs"][log_name]
        p = read_donnees(os.path.join(golden, "divergent_canal.donnees.f"))
        p.debitglob[0] = float(log["debitglob_1"])
        p.h_sortie[0] = float(log["h_sortie_1"])
        p.override_manning = n
        p.ts = 1e300
        U0 = initial_state(p, m, prep.zm)
        nsteps = log["series"][-1][0]
        print(f"{log_name}: {nsteps} steps", flush=True)
        ref = O.run(prep, p, U0, nsteps)
        with Solver(m, p) as s:
            s.set_discharge(True)
            s.set_state(U0)
            hist, done, st = s.step(nsteps)
            U = s.get_state()
            q_in, q_out, steady = s.discharge()
        assert st == 0 and done == nsteps and not steady
        _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)
        # the reference's printed outflow discharges, from the GPU run
        for it, _, qi, _, qo in log["series"]:
            assert abs(q_out[it - 1] - qo) <= 4 * np.spacing(qo), (log_name, it, q_out[it - 1], qo)
            assert abs(q_in[it - 1] - qi) <= 4 * np.spacing(qi), (log_name, it, q_in[it - 1], qi)


@pytest.mark.parametrize("options", [0, 2, 8 | 16])
def test_discharge_and_steady_state(options):
    """Per-step inflow/outflow discharges (summed in the caller's cell
    order) equal the oracle's bit for bit, over chunked runs; with
    tol_reg_perm the run stops at the oracle's step with the same tc, and
    stays stopped until a new state is set."""
    import json
    import os

    from cuteflow_amd import mesh as Mm
    from cuteflow_amd.params import read_donnees
    from cuteflow_amd.solver import initial_state

    m, p, U0 = river_case(nx=60, ny=12, q_peak=400.0, t_ramp=20.0)
    ref = O.run(O.Prepared(m), p, U0, 500, qout=True)
    with Solver(m, p, options=options) as s:
        s.set_discharge(True)
        s.set_state(U0)
        parts = []
        for k in (120, 380):
            s.step(k)
            parts.append(s.discharge())
        q_in = np.concatenate([x[0] for x in parts])
        q_out = np.concatenate([x[1] for x in parts])
    assert np.array_equal(q_in.view(np.int64), ref["qin"].view(np.int64))
    assert np.array_equal(q_out.view(np.int64), ref["qout"].view(np.int64))
    assert np.any(q_out > 0) and q_in[0] == 0.0 and np.all(q_in[1:] > 0)  # the hydrograph starts at 0

    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    mc = Mm.read_text_mesh(os.path.join(golden, "divergent_canal.mesh"))
    prep = O.Prepared(mc)
    pc = read_donnees(os.path.join(golden, "divergent_canal.donnees.f"))
    pc.debitglob[0], pc.h_sortie[0], pc.override_manning, pc.ts = 0.5, 0.02, 0.05, 1e300
    pc.tol_reg_perm = 1e-9
    Uc = initial_state(pc, mc, prep.zm)
    refc = O.run(prep, pc, Uc, 9000, qout=True)
    assert refc["steady"] and refc["steps"] < 9000
    with Solver(mc, pc, options=options) as s:
        s.set_state(Uc)
        hist, done, st = s.step(9000)
        q_in, q_out, steady = s.discharge()
        assert st == 0 and steady and done == refc["steps"] and s.tc == refc["tc"]
        _compare(s.get_state(), refc["U"], hist, refc["dt_hist"], pc.tolisec)
        assert np.array_equal(q_out, refc["qout"]) and np.array_equal(q_in, refc["qin"])
        assert s.step(10)[1] == 0           # stopped until a new state
        s.set_state(Uc)
        assert s.step(10)[1] == 10


@pytest.mark.parametrize("ilim", [0, 1, 2, 3, 4])
def test_muscl_limiters(ilim):
    """iflux 4 (solver_Riemann_muscl) on the wetting reach with semi-implicit
    friction, every slope limiter (4 is unset in the reference: first order)."""
    m, p, U0 = river_case(iflux=4, ilimiteur=ilim, nx=30, ny=9)
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 150)
    assert st == 0 and ref["status"] == 0 and done == 150
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


@pytest.mark.parametrize("fricimplic", [0, 2])
def test_muscl_friction_modes(fricimplic):
    if fricimplic == 0:
        m, p, U0 = dam_case(24)
        p.iflux, p.ilimiteur, p.friction, p.fricimplic = 4, 3, 1, 0
        m.manning[:] = 0.03
    else:
        m, p, U0 = river_case(iflux=4, ilimiteur=1, fricimplic=fricimplic)
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 300)
    assert st == 0 and ref["status"] == 0 and done == 300
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


@pytest.mark.parametrize("n", [20, 60])
def test_dam_break_muscl(n):
    m, p, U0 = dam_case(n)
    p.iflux, p.ilimiteur = 4, 1
    ref, U, hist, done, st, _, tc = _run_both(m, p, U0, 400)
    assert st == 0 and done == 400
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


@pytest.mark.parametrize("options", [2, 8 | 16, 4 | 1 | 16])
def test_muscl_layout_options_identical(options):
    """MUSCL under the tile sizes, dry-tile skipping and tile orders."""
    m, p, U0 = river_case(nx=120, ny=24, q_peak=2000.0, t_ramp=30.0, iflux=4, ilimiteur=3)
    prep = O.Prepared(m)
    ref = O.run(prep, p, U0, 600)
    with Solver(m, p, options=options) as s:
        s.set_state(U0)
        hist, done, st = s.step(600)
        U = s.get_state()
    assert st == 0 and done == 600
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


@pytest.mark.parametrize("n", [20, 60])
def test_dam_break_riadh(n):
    m, p, U0 = dam_case(n)
    p.iflux = 2
    ref, U, hist, done, st, _, tc = _run_both(m, p, U0, 400)
    assert st == 0 and done == 400
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


def test_river_hllc_transmissive_outlet():
    m, p, U0 = river_case(iflux=1, outlet="transm")
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 300)
    assert st == 0 and done == 300
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


@pytest.mark.parametrize("options", [0, 1, 2, 3, 4, 8, 4 | 1, 8 | 2, 16, 16 | 8, 16 | 4 | 2, 16 | 1])
def test_layout_options_identical(options):
    """Internal Hilbert order, dry-tile skipping (on/off), the tile size
    (256, 128, 64 cells) and the workgroup -> tile order (plain or
    XCD-contiguous) give the oracle's bits on a wetting floodplain over many
    steps."""
    m, p, U0 = river_case(nx=120, ny=24, q_peak=2000.0, t_ramp=30.0)
    prep = O.Prepared(m)
    ref = O.run(prep, p, U0, 800)
    with Solver(m, p, options=options) as s:
        s.set_state(U0)
        hist, done, st = s.step(400)
        hist2, done2, st2 = s.step(400)
        U = s.get_state()
        stats = s.last_run_stats()
    assert st == st2 == 0 and done == done2 == 400
    _compare(U, ref["U"], np.concatenate([hist[:-1], hist2]), ref["dt_hist"], p.tolisec)
    if options & 2:
        assert stats["active_fraction"] == 1.0
    tile = 64 if options & 8 else (128 if options & 4 else 256)
    assert stats["tile_cells"] == tile
    assert stats["tiles_total"] == 400 * ((m.nelt + tile - 1) // tile)


def test_reordering_invariance():
    """RCM and random cell orders give the same per-cell results."""
    m, p, U0 = river_case(nx=40, ny=10)
    for perm in (M.rcm_order(m), M.random_order(m, 7)):
        ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 200, perm=perm)
        _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


@pytest.mark.parametrize("iflux", [1, 2, 3, 4])
def test_lake_at_rest(iflux):
    m, p = lake_case(iflux=iflux)
    prep = O.Prepared(m)
    U0 = np.zeros((m.nelt, 3))
    U0[:, 0] = np.where(-prep.zm > p.tolisec, -prep.zm, 0.0)
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 200)
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)
    wet = U[:, 0] > 0
    assert np.max(np.abs(U[wet, 0] + prep.zm[wet])) < 1e-6
    assert np.max(np.abs(U[:, 1:])) < 1e-6


def test_stop_at_ts():
    m, p, U0 = dam_case(16)
    p.ts = 0.5
    prep = O.Prepared(m)
    ref = O.run(prep, p, U0, 10000)
    with Solver(m, p) as s:
        s.set_state(U0)
        hist, done, st = s.step(10000)
        U = s.get_state()
        assert done == ref["steps"] < 10000
        assert s.tc == ref["tc"]
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


def test_dt_error_all_dry():
    """An all-dry domain gives dt = 1e5 > 9999: the run stops with an error
    after the first step (cfl_condition.cuf:129-132)."""
    m, p, U0 = dam_case(10)
    U0[:] = 0.0
    with Solver(m, p) as s:
        s.set_state(U0)
        hist, done, st = s.step(10, raise_on_error=False)
        assert st == -3 and done == 1
        with pytest.raises(RuntimeError):
            s.step(1)


def test_nan_detected():
    m, p, U0 = dam_case(10)
    U0[5, 1] = np.nan
    with Solver(m, p) as s:
        s.set_state(U0)
        hist, done, st = s.step(10, raise_on_error=False)
        assert st == -3 and done == 1


def test_constant_dt():
    m, p, U0 = dam_case(12)
    p.is_dt_constant, p.constant_dt = 1, 0.01
    ref, U, hist, done, st, _, _ = _run_both(m, p, U0, 50)
    assert np.all(hist == 0.01)
    _compare(U, ref["U"], hist, ref["dt_hist"], p.tolisec)


def test_chunked_steps_equal_single_run():
    m, p, U0 = river_case(nx=30, ny=9)
    with Solver(m, p) as s:
        s.set_state(U0)
        h1, _, _ = s.step(120)
        U1 = s.get_state()
    with Solver(m, p) as s:
        s.set_state(U0)
        parts = [s.step(k)[0][:-1] for k in (1, 7, 50, 62)]
        U2 = s.get_state()
    assert np.array_equal(np.concatenate(parts), h1[:-1])
    assert np.array_equal(U1, U2)


def _tiny_case(nx, ny, iflux, seed):
    """nx x ny quads (2 nx ny triangles) on a sloped, jittered bed with walls
    all around, per-cell n, and a dam state: a wet block of two depths next
    to dry cells."""
    rng = np.random.default_rng(seed)
    x, y = M._grid_nodes(nx, ny, 10.0 * nx, 10.0 * ny, 0.2, rng)
    z = -0.01 * x + 0.05 * rng.standard_normal(x.shape)
    conn = M._grid_triangles(nx, ny)
    _, walls = M._boundary_nodes(nx, ny)
    m = M.Mesh(coords=np.stack([x, y, z], 1), conn=conn, manning=rng.uniform(0.02, 0.06, len(conn)), ndwall=walls)
    p = Params(iflux=iflux, ilimiteur=1, tolisec=1e-3, cfl=0.9, ts=1e300, inlet="transm", outlet="transm",
                 friction=1, fricimplic=1)
    xc, _ = m.centroids()
    U0 = np.zeros((m.nelt, 3))
    U0[:, 0] = np.where(xc < 3.0 * nx, 2.0, np.where(xc < 7.0 * nx, 0.5, 0.0))
    U0[:, 1] = np.where(U0[:, 0] > 0, 0.1 * U0[:, 0], 0.0)
    return m, p, U0


@pytest.mark.parametrize("nx,ny", [(1, 1), (2, 2), (4, 8), (33, 1), (129, 1)])
@pytest.mark.parametrize("iflux", [1, 2, 3, 4])
def test_tiny_meshes_at_tile_edges(nx, ny, iflux):
    """Meshes of 2, 8, 64, 66 and 258 triangles (one partial tile, exactly one
    64-cell tile, one tile and two cells, one 256-cell tile and two cells),
    every scheme, 256- and 64-cell tiles: the same bits as the oracle."""
    from cuteflow_amd import _lib
    m, p, U0 = _tiny_case(nx, ny, iflux, seed=nx * 7 + ny)
    nsteps = 30
    ref = O.run(O.Prepared(m), p, U0, nsteps)
    for options in (0, _lib.CF_OPT_TILE_64 | _lib.CF_OPT_XCD_CONTIG):
        with Solver(m, p, options=options) as s:
            s.set_state(U0)
            hist, done, st = s.step(nsteps, raise_on_error=False)
            U = s.get_state()
        assert st == ref["status"] and done == ref["steps"], (options, st, done)
        _compare(U, ref["U"], np.asarray(hist), ref["dt_hist"], p.tolisec)


@pytest.mark.parametrize("iflux", [1, 2, 3, 4])
@pytest.mark.parametrize("options", [0, 24])
def test_phys_inflow_matches_oracle(iflux, options):
    """inlet = 'phys_inflow' (the inlet discharge shared in proportion to the
    depth, flux_riemann.cuf:127-137, the wetted inlet area recomputed before
    every step): states, dt and the per-step discharges equal the oracle's
    bit for bit, alone and next to an 'inflow' instance in a batch; the
    adjoint refuses it."""
    from cuteflow_amd import _lib
    from cuteflow_amd.adjoint import AdjointRun
    from cuteflow_amd.solver import BatchSolver
    m, p, U0 = river_case(nx=60, ny=12, seed=32, iflux=iflux, q_peak=400.0, t_ramp=20.0)
    p.inlet = "phys_inflow"
    nsteps = 300
    ref = O.run(O.Prepared(m), p, U0, nsteps, qout=True)
    assert ref["status"] == 0 and np.all(ref["qin"][1:] > 0)
    with Solver(m, p, options=options) as s:
        s.set_discharge(True)
        s.set_state(U0)
        hist, done, st = s.step(nsteps)
        q_in, q_out, _ = s.discharge()
        assert st == 0 and done == nsteps
        _compare(s.get_state(), ref["U"], np.asarray(hist), ref["dt_hist"], p.tolisec)
        assert np.array_equal(q_in.view(np.int64), ref["qin"].view(np.int64))
        assert np.array_equal(q_out.view(np.int64), ref["qout"].view(np.int64))
        if options == 0:
            run = AdjointRun(s, [np.array([0], np.int32)], sample_every=10, checkpoint_every=10)
            with pytest.raises(_lib.CuteflowError) as err:
                run.forward(20, U0=[U0])
            assert err.value.status == _lib.CF_ERR_ARG
    m2, p2, U2 = river_case(nx=40, ny=10, seed=33, iflux=iflux)
    ref2 = O.run(O.Prepared(m2), p2, U2, nsteps)
    with BatchSolver([m, m2], [p, p2], options=options) as bs:
        bs.set_state(0, U0)
        bs.set_state(1, U2)
        hist, done, status = bs.step(nsteps)
        for b, r in enumerate((ref, ref2)):
            assert done[b] == r["steps"]
            assert np.array_equal(np.asarray(hist[b]).view(np.int64), r["dt_hist"].view(np.int64))
            assert np.array_equal(bs.get_state(b).view(np.int64), r["U"].view(np.int64))
