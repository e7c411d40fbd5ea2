#!/usr/bin/env python3
"""Benchmark of the FP64 finite-volume St-Venant time step on one MI355X.

Headline (value, roofline, cpu_baseline): the river-reach configuration,
a 20 km x 2 km reach meshed as 3200 x 320 jittered quads split into
2,048,000 triangles in RCM order, bed slope 1/2000 + 0.5 m sinusoidal bars
(300 m) + a seeded Gaussian-smoothed random field (sigma 0.3 m, 200 m
correlation), a 3 m deep main channel in a floodplain that starts dry,
Manning n 0.03 / 0.06, inflow ramped 0 -> 3000 m3/s over 1 h, fixed outlet
level, dry threshold 1e-3 m, WAF second order with the superbee limiter and
semi-implicit friction (fricimplic 1).  A step is one time step over the
whole mesh.  The schedule (cells per tile, dry-tile skipping, tile order
over the XCDs) is picked by the autotuner before the run (--options auto),
outside the timed region.

Further sections of the same JSON line (each skippable):
- "c1": the 20,000-cell circular dam break (HLLC, walls) for its full 1000
  steps: GPU time per step, and the oracle's whole run on all host threads
  and on one, with the final states compared bit for bit;
- "calibration": adjoint/forward cost ratio of a gauge-misfit gradient on a
  100k-cell reach with 8 Manning zones, 20 gauges sampled every 10 steps
  over 2000 steps, checkpoints every 100 steps; the gradient checked
  against central differences and tangent-linear derivatives on each zone;
  20 L-BFGS iterations from n 20 % above the truth;
- "ensemble": 32 members of 50k-500k cells in one launch per step, 2000
  steps, with the schedule (cells per tile, dry-tile skipping, tile order
  over the XCDs) picked by the autotuner, against the default schedule and
  against the same 32 runs one after the other;
- "batched_gradients": 16 members on a 1M-cell reach with per-cell
  log-normal Manning fields and their own hydrographs, 3000 steps,
  checkpoints every 150 steps, loss = misfit at 50 gauges + squared error of
  the maximum flood extent; gradients per second, cost ratio to the batched
  forward, and tangent-linear checks along 4 random directions.
The gradient workloads use the first-order scheme (iflux 1), as the
reference's own calibration runs do; see the "scheme" fields.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic HBM bytes per cell-update of the WAF step with per-cell n:
# state in 24 + out 24, neighbour ids 12, vertex x, y 48, zm/surf/cotemin 24,
# n 8, centroid x, y 16
BYTES_WAF = 156
BYTES_HLLC = 132
PEAK_HBM_GBS = 8000.0
METRIC = "FP64 cell-updates/s + %HBM roofline; adjoint/forward cost ratio; batched grads/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(nx, ny, order, seed=1):
    from cuteflow_amd import scenarios as SC

    return SC.river_reach(nx, ny, order, seed)


def traffic_from_profile(workload_key, options):
    """HBM bytes per step-kernel launch from the committed counter profile of
    this workload, if it was profiled with the same schedule options (the
    traffic depends on the tile size, order and skipping), else None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_key}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        prof = json.load(fh)
    if prof.get("schedule_options") != options:
        return None
    return prof.get("hbm_bytes_per_launch")


def sync():
    import torch

    torch.cuda.synchronize()


def cpu_info():
    """Host CPU model and the cores this process may use (affinity mask and
    the cgroup CPU quota, when one is set)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, per = fh.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    usable = len(os.sched_getaffinity(0))
    if quota is not None:
        usable = max(1, min(usable, int(quota)))
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "cgroup_cpu_quota": quota, "usable_cores": usable}


def timed_run(s, nsteps):
    """nsteps steps of solver s, already prepared for that step count:
    returns (wall seconds between two device synchronisations, seconds on
    the solver's stream by HIP events, run statistics)."""
    import ctypes as C

    import numpy as np
    import torch

    from cuteflow_amd import _lib

    stream = torch.cuda.ExternalStream(s.stream)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    hist = np.zeros(nsteps + 1)
    done = C.c_int()
    sync()
    w0 = time.perf_counter()
    ev0.record(stream)
    _lib.check(s.lib.cf_step_async(s.h, nsteps), "cf_step_async")
    ev1.record(stream)
    _lib.check(s.lib.cf_step_finish(s.h, _lib.dptr(hist), C.byref(done)), "cf_step_finish")
    sync()
    w1 = time.perf_counter()
    if done.value != nsteps:
        raise RuntimeError(f"run stopped after {done.value} of {nsteps} steps")
    return w1 - w0, ev0.elapsed_time(ev1) / 1000.0, s.last_run_stats()


# ---------------------------------------------------------------------------
def bench_headline(args):
    import numpy as np
    import torch

    from cuteflow_amd.solver import Solver, autotune

    t0 = time.perf_counter()
    m, p, U0, perm = make_workload(args.nx, args.ny, args.order)
    log(f"bench: workload nelt={m.nelt} order={args.order} built in {time.perf_counter() - t0:.1f}s")
    schedule = {"autotuned": args.options == "auto"}
    if args.options == "auto":
        t0 = time.perf_counter()
        options, table = autotune([m], [p], [U0], nsteps=args.tune_steps)
        schedule.update({"tune_steps": args.tune_steps, "tune_seconds": time.perf_counter() - t0,
                         "candidates": [dict(describe_options(o), options=o, ms=ms) for o, ms in table]})
        log(f"bench: autotune {schedule['tune_seconds']:.1f} s, best {describe_options(options)}")
    else:
        options = int(args.options)
    schedule.update(describe_options(options), options=options)
    s = Solver(m, p, options=options)
    s.set_state(U0)
    if args.warmup:
        s.step(args.warmup)
    # one-time work of a run of args.steps steps (history, graph capture)
    # stays out of the timed region
    s.prepare(args.steps)
    log(f"bench: warmup {args.warmup} steps done, tc={s.tc:.3f}s")

    if args.copy_probe_mb:
        n = args.copy_probe_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        sync()
        log(f"bench: copy probe {args.copy_probe_mb} MiB read + {args.copy_probe_mb} MiB written, 3 times")
        del src, dst

    wall, t_stream, stats = timed_run(s, args.steps)
    per_launch_s = t_stream / args.steps
    value = m.nelt * args.steps / wall
    # cells the step kernel actually processed per launch: dry tiles with dry
    # neighbours are skipped (their update is the identity)
    active_cells = stats["active_fraction"] * m.nelt
    achieved = BYTES_WAF * active_cells / per_launch_s / 1e9
    U_end = s.get_state()
    wet = int(np.sum(U_end[:, 0] > p.tolisec))
    log(f"bench: {args.steps} steps in {wall * 1e3:.3f} ms wall, {t_stream * 1e3:.3f} ms on the stream, "
        f"{value:.3e} cell-updates/s, active tiles {stats['active_fraction']:.3f}, "
        f"{achieved:.0f} GB/s algorithmic over processed cells, wet cells {wet}")

    workload_key = f"river_waf_{m.nelt}_{args.order}"
    traffic = traffic_from_profile(workload_key, options)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, no external data)",
        "config": {
            "workload": "river reach with floodplain: WAF 2nd order + semi-implicit Manning friction, "
                        f"{args.nx}x{args.ny}x2 jittered triangles, {args.order} order, inflow ramp "
                        "0->3000 m3/s over 1 h, fixed outlet level, dry threshold 1e-3 m",
            "nelt": m.nelt,
            "order": args.order,
            "device_order": "caller's" if options & 1 else "Hilbert curve of centroids",
            "dry_tile_skipping": not (options & 2),
            "schedule": schedule,
            "iflux": 3,
            "fricimplic": 1,
            "tc_end": s.tc,
            "wet_cells_end": wet,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "kernel": "step_kernel<3,2,0>",
            "bytes_per_cell_update": BYTES_WAF,
            "cells_per_launch": active_cells,
            "active_tile_fraction": stats["active_fraction"],
            "avg_launch_us": per_launch_s * 1e6,
            "timing": "HIP events on the solver stream over the timed steps (step kernel + "
                      "one-workgroup finalize kernel per step)",
        },
        "processed_cell_updates_per_s": active_cells * args.steps / wall,
        "processed_cell_updates_definition": "cells of the tiles the step kernel ran (wet tiles and their "
                                             "neighbours) x steps / wall time; value counts every cell",
    }
    workload_desc = result["config"]["workload"]

    if not args.extra_runs:
        if args.cpu_seconds > 0:
            raise SystemExit("--extra-runs 0 is for profiling the timed steps: use it with --cpu-seconds 0")
        s.close()
        return result

    # The same run with every tile computed every step (no dry-tile
    # skipping): the cost of the whole mesh, and the bits of the skipping run.
    from cuteflow_amd import _lib
    with Solver(m, p, options=options | _lib.CF_OPT_NO_TILE_SKIP) as sa:
        sa.set_state(U0)
        if args.warmup:
            sa.step(args.warmup)
        sa.prepare(args.steps)
        wall_a, ts_a, _ = timed_run(sa, args.steps)
        same_a = bool(np.array_equal(sa.get_state(), U_end))
    ach_a = BYTES_WAF * m.nelt / (ts_a / args.steps) / 1e9
    log(f"bench: every tile: {wall_a * 1e6 / args.steps:.1f} us/step, {ach_a:.0f} GB/s algorithmic, "
        f"bit-identical to the skipping run: {same_a}")
    result["all_tiles"] = {
        "workload": workload_desc + ", every tile computed every step",
        "ms_per_step": wall_a * 1000.0 / args.steps, "cell_updates_per_s": m.nelt * args.steps / wall_a,
        "avg_launch_us": ts_a / args.steps * 1e6, "achieved_gbs": ach_a, "frac": ach_a / PEAK_HBM_GBS,
        "bit_identical_to_skipping_run": same_a}

    # The same reach stored in random cell order (the cells are laid out on
    # the device along a Hilbert curve whatever the input order).
    if args.order != "random":
        mr, pr, U0r, perm_r = make_workload(args.nx, args.ny, "random")
        with Solver(mr, pr, options=options) as sr:
            sr.set_state(U0r)
            if args.warmup:
                sr.step(args.warmup)
            sr.prepare(args.steps)
            wall_r, ts_r, st_r = timed_run(sr, args.steps)
            Ur = sr.get_state()
        nat_a, nat_r = np.empty_like(U_end), np.empty_like(Ur)
        nat_a[perm], nat_r[perm_r] = U_end, Ur
        same_r = bool(np.array_equal(nat_a, nat_r))
        log(f"bench: random order: {wall_r * 1e6 / args.steps:.1f} us/step, bit-identical to the "
            f"{args.order}-order run: {same_r}")
        result["random_order"] = {
            "workload": workload_desc.replace(f"{args.order} order", "random order"),
            "ms_per_step": wall_r * 1000.0 / args.steps, "cell_updates_per_s": m.nelt * args.steps / wall_r,
            "avg_launch_us": ts_r / args.steps * 1e6, "active_tile_fraction": st_r["active_fraction"],
            "bit_identical_to_headline": same_r}

    # CPU baseline: the oracle on a bounded sample of the same workload, on
    # the host cores, started from the same initial state; the GPU result
    # after the same number of steps is compared bit for bit.
    if args.cpu_seconds > 0:
        from oracle import oracle as O

        host = cpu_info()
        threads = args.cpu_threads or host["usable_cores"]
        t0 = time.perf_counter()
        prep = O.Prepared(m)
        log(f"bench: oracle preprocessing {time.perf_counter() - t0:.1f}s; host {host}")
        t0 = time.perf_counter()
        O.run(prep, p, U0, 2, nthreads=threads)
        per_step = (time.perf_counter() - t0) / 2
        nsamp = max(2, min(2000, int(args.cpu_seconds / max(per_step, 1e-6))))
        t0 = time.perf_counter()
        ref = O.run(prep, p, U0, nsamp, nthreads=threads)
        cpu_t = time.perf_counter() - t0
        s.set_state(U0)
        hist_g, done_g, _ = s.step(nsamp)
        Ug = s.get_state()
        same = bool(np.array_equal(Ug, ref["U"]) and np.array_equal(hist_g, ref["dt_hist"]))
        log(f"bench: CPU oracle {nsamp} steps in {cpu_t:.1f}s on {threads} threads; GPU bit-identical: {same}")
        result["cpu_baseline"] = {
            "value": m.nelt * nsamp / cpu_t,
            "unit": "cell-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"first {nsamp} steps of the same workload from the same initial state "
                      f"(C restatement of the reference kernels, gcc -O3, OpenMP, {threads} threads = every "
                      f"core this process may use, 1 MPI rank)",
            "cpu": host,
        }
        # the same oracle on one host thread, over a shorter sample: the
        # single-core rate beside the all-core one
        t0 = time.perf_counter()
        O.run(prep, p, U0, 1, nthreads=1)
        per_step1 = time.perf_counter() - t0
        n1 = max(1, min(nsamp, int(args.cpu_seconds / 4.0 / max(per_step1, 1e-6))))
        t0 = time.perf_counter()
        O.run(prep, p, U0, n1, nthreads=1)
        cpu_t1 = time.perf_counter() - t0
        log(f"bench: CPU oracle {n1} steps in {cpu_t1:.1f}s on 1 thread")
        result["cpu_baseline"]["single_core"] = {"value": m.nelt * n1 / cpu_t1, "unit": "cell-updates/s",
                                                 "cores": 1, "sample": f"first {n1} steps, 1 thread"}
        result["parity_check"] = {"steps": nsamp, "bit_identical": same}
    s.close()
    return result


# ---------------------------------------------------------------------------
def bench_c1(args):
    """The small configuration in full: circular dam break on a 200 m x 200 m
    square of 100 x 100 jittered quads (20,000 triangles), HLLC, walls, no
    friction, 1000 steps; GPU time per step, the oracle's whole run on all
    host threads and on one, and the bit comparison of the final states."""
    import ctypes as C

    import numpy as np
    import torch

    from cuteflow_amd import _lib
    from cuteflow_amd import mesh as M
    from cuteflow_amd.params import Params
    from cuteflow_amd.solver import Solver
    from oracle import oracle as O

    nsteps = 1000
    m = M.dam_break_mesh(100, size=200.0, jitter=0.2, seed=0)
    p = Params(iflux=1, tolisec=1e-5, cfl=0.9, ts=1e300, inlet="transm", outlet="transm", friction=0)
    U0 = M.circular_dam_state(m)
    out = {"workload": "circular dam break, 100 x 100 jittered quads = 20,000 triangles, HLLC, walls, no friction, "
                       "1000 steps (h 10 m inside r = 30 m, 1 m outside)", "nelt": m.nelt, "steps": nsteps}
    options, out["schedule"] = section_options(args, [m], [p], [U0], "c1")
    with Solver(m, p, options=options) as s:
        stream = torch.cuda.ExternalStream(s.stream)
        hist = np.zeros(nsteps + 1)
        done = C.c_int()
        times = []
        for _ in range(2):  # the first run captures the step graph
            s.set_state(U0)
            sync()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            _lib.check(s.lib.cf_step_async(s.h, nsteps), "cf_step_async")
            ev1.record(stream)
            _lib.check(s.lib.cf_step_finish(s.h, _lib.dptr(hist), C.byref(done)), "cf_step_finish")
            sync()
            times.append(ev0.elapsed_time(ev1) / 1000.0)
        Ug = s.get_state()
    t_gpu = times[-1]
    out.update({"gpu_us_per_step": 1e6 * t_gpu / nsteps, "gpu_cell_updates_per_s": m.nelt * nsteps / t_gpu})
    log(f"bench: c1 GPU {1e6 * t_gpu / nsteps:.2f} us/step, {m.nelt * nsteps / t_gpu:.3e} cell-updates/s")
    if args.cpu_seconds > 0:
        prep = O.Prepared(m)
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        for nt in (threads, 1):
            t = time.perf_counter()
            ref = O.run(prep, p, U0, nsteps, nthreads=nt)
            dt = time.perf_counter() - t
            same = bool(np.array_equal(Ug, ref["U"]) and np.array_equal(hist[: nsteps + 1], ref["dt_hist"]))
            out[f"cpu_{'all' if nt == threads else 'one'}_core"] = {
                "threads": nt, "seconds": dt, "cell_updates_per_s": m.nelt * nsteps / dt, "gpu_bit_identical": same}
            log(f"bench: c1 CPU oracle {nsteps} steps in {dt:.2f}s on {nt} threads; GPU bit-identical: {same}")
    return out


def bench_calibration(args):
    """Zonal calibration reach, with each first-order scheme: HLLC (iflux 1)
    and the reference's calibration scheme (iflux 2)."""
    out = {"workload": "100k-cell reach starting flooded (surface 0.5 m above the mean floodplain level, "
                       "inclined like the bed), 8 Manning zones (4 reaches x channel/floodplain), 20 gauges "
                       "(2-3 per zone), levels every 10 steps over 2000 steps, checkpoints every 100 steps, "
                       "semi-implicit friction; observations from the hidden truth, gradient at +20 %"}
    for iflux in (1, 2):
        out[f"iflux{iflux}"] = _calibration_one(args, iflux)
    return out


def _calibration_one(args, iflux):
    """Zonal calibration reach: gradient cost, checks, L-BFGS."""
    import numpy as np

    from cuteflow_amd import calibration as CAL
    from cuteflow_amd import scenarios as SC
    from cuteflow_amd.adjoint import AdjointRun
    from cuteflow_amd.solver import Solver

    nsteps, every, K = 2000, 10, 100
    sc = SC.calibration(100_000, ngauges=20, iflux=iflux)
    N = sc.mesh.nelt
    Z = np.zeros((N, 8))
    Z[np.arange(N), sc.zone] = 1.0
    out = {"nelt": N, "scheme": f"iflux {iflux} ({'HLLC' if iflux == 1 else 'Riadh state HLLC'}), fricimplic 1"}
    options, out["schedule"] = section_options(args, [sc.mesh], [sc.params], [sc.state], f"calibration iflux {iflux}")
    with Solver(sc.mesh, sc.params, options=options) as s:
        s.set_state(sc.state)
        s.step(nsteps)
        times = []
        for _ in range(2):
            s.set_state(sc.state)
            sync()
            t = time.perf_counter()
            s.step(nsteps)
            sync()
            times.append(time.perf_counter() - t)
        t_fwd = min(times)
        run = AdjointRun(s, sc.gauges, sample_every=every, checkpoint_every=K)
        obs = run.forward(nsteps, U0=[sc.state], manning=[Z @ sc.n_zones])[0][0]
        n0 = sc.n_zones * 1.2
        grads, gtimes = [], []
        for _ in range(3):
            sync()
            t = time.perf_counter()
            eta = run.forward(nsteps, U0=[sc.state], manning=[Z @ n0])[0][0]
            gn, _ = run.backward([2.0 * (eta - obs)])
            sync()
            gtimes.append(time.perf_counter() - t)
            grads.append(Z.T @ gn[0])
        t_grad = min(gtimes[1:])
        g = grads[-1]
        reproducible = all(np.array_equal(grads[0], x) for x in grads)
        log(f"bench: calibration iflux {iflux} forward {t_fwd*1e3:.1f} ms, gradient {t_grad*1e3:.1f} ms "
            f"(first {gtimes[0]*1e3:.1f} ms), ratio {t_grad / t_fwd:.2f}")
        g_tlm = CAL.zonal_tlm_gradient(run, Z, obs)
        g_fd, cf6, bf6 = CAL.zonal_fd_gradient(run, Z, n0, obs, eps=1e-6, count_flips=True)
        g_fd7, cf7, bf7 = CAL.zonal_fd_gradient(run, Z, n0, obs, eps=1e-7, count_flips=True)
        rel = lambda a, b: (np.abs(a - b) / np.maximum(np.abs(b), 1e-300)).tolist()
        log(f"bench: calibration iflux {iflux} gradient vs TLM {max(rel(g, g_tlm)):.1e}, vs FD(1e-7) "
            f"{rel(g, g_fd7)} with branch flips {bf7.tolist()}")
        out.update({
            "forward_ms": t_fwd * 1e3,
            "gradient_ms": t_grad * 1e3,
            "gradient_first_call_ms": gtimes[0] * 1e3,
            "adjoint_forward_ratio": t_grad / t_fwd,
            "gradient_ratio_definition": "(recorded forward + reverse sweep with checkpoint recompute) / plain "
                                         "forward of the same 2000 steps, steady state",
            "speedup_vs_fd_zonal": 16.0 * t_fwd / t_grad,
            "speedup_vs_fd_per_cell": 2.0 * N * t_fwd / t_grad,
            "fd_cost_definition": "central differences need 2 plain forwards per parameter (8 zones: 16; "
                                  "per-cell n: 2 x nelt)",
            "gradient": g.tolist(),
            "rel_err_vs_tlm": rel(g, g_tlm),
            "rel_err_vs_fd_1e-6": rel(g, g_fd),
            "rel_err_vs_fd_1e-7": rel(g, g_fd7),
            "fd_clamp_flips_1e-6": cf6.tolist(), "fd_branch_flips_1e-6": bf6.tolist(),
            "fd_clamp_flips_1e-7": cf7.tolist(), "fd_branch_flips_1e-7": bf7.tolist(),
            "fd_flips_definition": "per zone: (step, cell) updates whose wet/dry clamp (clamp) or any branch "
                                   "(branch) the two perturbed runs take differently from the run at n0",
            "fd_within_1e-6_on_branch_flip_free_zones": bool(all(
                e <= 1e-6 for e, f in zip(rel(g, g_fd7) + rel(g, g_fd), bf7.tolist() + bf6.tolist()) if f == 0)),
            "reproducible_bits": bool(reproducible),
        })
        if args.lbfgs_iters > 0:
            t = time.perf_counter()
            nz, hist, nev = CAL.zonal_lbfgs(run, sc.zone, n0, obs, nsteps, sc.state, iters=args.lbfgs_iters)
            tl = time.perf_counter() - t
            log(f"bench: L-BFGS {args.lbfgs_iters} iterations ({nev} evaluations) in {tl:.1f}s, "
                f"misfit {hist[0]:.3e} -> {min(hist):.3e}")
            out["lbfgs"] = {"iterations": args.lbfgs_iters, "evaluations": nev, "seconds": tl,
                            "relative_misfit_start": hist[0], "relative_misfit_best": min(hist),
                            "n_truth": sc.n_zones.tolist(), "n_start": n0.tolist(), "n_final": nz.tolist(),
                            "max_rel_err_n_final": float(np.max(np.abs(nz - sc.n_zones) / sc.n_zones))}
    return out


# ---------------------------------------------------------------------------
def describe_options(o):
    """Schedule named by cf_params.options bits."""
    from cuteflow_amd import _lib
    return {"tile_cells": 64 if o & _lib.CF_OPT_TILE_64 else (128 if o & _lib.CF_OPT_TILE_128 else 256),
            "skip_dry_tiles": not (o & _lib.CF_OPT_NO_TILE_SKIP), "xcd_contiguous": bool(o & _lib.CF_OPT_XCD_CONTIG)}


def section_options(args, meshes, params, states, label):
    """Schedule of a section's runs: the autotuner's pick (--options auto)
    or the given options bits; returns (options, record)."""
    if args.options != "auto":
        return int(args.options), {"autotuned": False, "options": int(args.options)}
    from cuteflow_amd.solver import autotune
    t = time.perf_counter()
    best, _ = autotune(meshes, params, states, nsteps=args.tune_steps)
    dt = time.perf_counter() - t
    log(f"bench: {label} autotune {dt:.1f} s, best {describe_options(best)}")
    return best, dict(describe_options(best), autotuned=True, options=best, tune_seconds=dt)


def bench_ensemble(args):
    """32 uneven members, one launch per step with the autotuned schedule,
    vs the default schedule and vs one run after the other."""
    import numpy as np

    from cuteflow_amd import scenarios as SC
    from cuteflow_amd.solver import BatchSolver, Solver, autotune

    nsteps = 2000
    ens = SC.ensemble(32)
    cells = sum(e.mesh.nelt for e in ens)
    meshes, params, states = [e.mesh for e in ens], [e.params for e in ens], [e.state for e in ens]
    out = {"workload": "32 members of the 20 km x 2 km reach meshed at 50k-500k cells, peak inflow 500-4000 "
                       "m3/s, n in 8 zones 0.02-0.08, WAF + semi-implicit friction, 2000 steps, own dt each",
           "members": len(ens), "cells": cells, "sizes": [e.mesh.nelt for e in ens]}
    t = time.perf_counter()
    best, table = autotune(meshes, params, states, nsteps=args.tune_steps)
    t_tune = time.perf_counter() - t
    log(f"bench: autotune {t_tune:.1f} s, best {describe_options(best)}")
    out["autotune"] = {"steps": args.tune_steps, "seconds": t_tune, "best_options": best, "best": describe_options(best),
                       "candidates": [dict(describe_options(o), options=o, ms=ms) for o, ms in table]}

    def batched(options):
        with BatchSolver(meshes, params, options=options) as bs:
            ts = []
            for _ in range(2):
                for b, U in enumerate(states):
                    bs.set_state(b, U)
                sync()
                t = time.perf_counter()
                hist, done, st = bs.step(nsteps)
                sync()
                ts.append(time.perf_counter() - t)
            assert int(done.min()) == nsteps and int(st.max()) == 0
            return ts[-1], bs.last_run_stats()["active_fraction"], [bs.get_state(b) for b in (0, len(ens) - 1)]

    t_batch, frac, u_best = batched(best)
    t_default, _, u_default = batched(0)
    same = all(np.array_equal(a, b) for a, b in zip(u_best, u_default))
    t_seq = 0.0
    for e in ens:
        with Solver(e.mesh, e.params) as s:
            for k in range(2):
                s.set_state(e.state)
                sync()
                t = time.perf_counter()
                s.step(nsteps)
                sync()
                if k == 1:
                    t_seq += time.perf_counter() - t
    log(f"bench: ensemble batched {t_batch*1e3:.0f} ms (default schedule {t_default*1e3:.0f} ms), "
        f"sequential {t_seq*1e3:.0f} ms, speedup {t_seq / t_batch:.2f}")
    out.update({"batched_s": t_batch, "batched_default_schedule_s": t_default, "sequential_s": t_seq,
                "speedup": t_seq / t_batch, "batched_cell_updates_per_s": cells * nsteps / t_batch,
                "sequential_cell_updates_per_s": cells * nsteps / t_seq, "active_tile_fraction": frac,
                "tuned_equals_default_bits": bool(same)})
    return out


# ---------------------------------------------------------------------------
def bench_field(args):
    """16 x 1M per-cell Manning fields: batched gradients."""
    import numpy as np
    import torch

    from cuteflow_amd import calibration as CAL
    from cuteflow_amd import scenarios as SC
    from cuteflow_amd.adjoint import AdjointRun
    from cuteflow_amd.solver import BatchSolver

    nsteps, every, K = 3000, 10, 150
    mem = SC.field_ensemble(16, 1_000_000, ngauges=50, iflux=1)
    B = len(mem)
    out = {"workload": "16 members on a 1M-cell 20 km x 2 km reach, per-cell log-normal n (median 0.04, "
                       "sigma_log 0.3, 500 m correlation), own hydrographs, 3000 steps, checkpoints every "
                       "150 steps, loss = misfit at 50 gauges (every 10 steps) + squared error of the soft "
                       "maximum flood extent; observations from the member's field, gradient at +20 %",
           "members": B, "cells": sum(e.mesh.nelt for e in mem), "scheme": "iflux 1 (HLLC), fricimplic 1"}
    options, out["schedule"] = section_options(args, [e.mesh for e in mem], [e.params for e in mem],
                                               [e.state for e in mem], "field ensemble")
    with BatchSolver([e.mesh for e in mem], [e.params for e in mem], options=options) as bs:
        U0 = [e.state for e in mem]
        surf = [bs.geometry(b)["surf"] for b in range(B)]
        # the loss and its seeds stay on the GPU (torch), as in a calibration loop
        surf_t = [torch.as_tensor(x, device="cuda") for x in surf]
        ts = []
        for _ in range(2):
            for b in range(B):
                bs.set_state(b, U0[b])
            sync()
            t = time.perf_counter()
            bs.step(nsteps)
            sync()
            ts.append(time.perf_counter() - t)
        t_fwd = ts[-1]
        run = AdjointRun(bs, [e.gauges for e in mem], sample_every=every, checkpoint_every=K, want_hmax=True)
        obs, hobs = run.forward(nsteps, U0=U0, manning=[e.mesh.manning for e in mem])
        a_obs = [CAL.soft_extent(hobs[b], surf[b])[0] for b in range(B)]
        n0 = [1.2 * e.mesh.manning for e in mem]
        scale = [max(1.0, 0.01 * a) for a in a_obs]
        # the initial states and the fields being calibrated are resident in
        # HBM, like the forward's initial states
        U0_d = [torch.as_tensor(u, device="cuda") for u in U0]
        n0_d = [torch.as_tensor(n, device="cuda") for n in n0]

        def gradient():
            etas, hm = run.forward(nsteps, U0=U0_d, manning=n0_d, device=True)
            ge, gh, loss = [], [], 0.0
            for b in range(B):
                ge.append(2.0 * (etas[b] - obs[b]))
                le, dh = CAL.extent_loss(hm[b], surf_t[b], a_obs[b], scale[b])
                gh.append(dh)
                loss += CAL.gauge_misfit(etas[b], obs[b]) + le
            gn, _ = run.backward(ge, gh, want_state=False, device=True)
            return gn, etas, hm, loss

        gt = []
        for _ in range(2):
            sync()
            t = time.perf_counter()
            gn, etas, hm, loss = gradient()
            sync()
            gt.append(time.perf_counter() - t)
        t_grad = gt[-1]
        gn = [x.cpu().numpy() for x in gn]
        hm = [x.cpu().numpy() for x in hm]
        log(f"bench: field ensemble forward {t_fwd*1e3:.0f} ms, gradient {t_grad*1e3:.0f} ms "
            f"(first {gt[0]*1e3:.0f} ms), ratio {t_grad / t_fwd:.2f}, {B / t_grad:.2f} grads/s")
        # tangent-linear checks along 4 random directions (per-cell smooth
        # relative perturbations of every member's field)
        rng = np.random.default_rng(3)
        checks = []
        for d in range(4):
            dn = [n0[b] * rng.normal(0.0, 0.1, n0[b].shape) for b in range(B)]
            tl = run.tlm(dn)
            dl_tlm = 0.0
            for b in range(B):
                dl_tlm += float(np.sum(2.0 * (etas[b] - obs[b]) * tl[b]["d_eta"]))
                _, dh = CAL.extent_loss(hm[b], surf[b], a_obs[b], scale[b])
                dl_tlm += float(np.sum(dh * tl[b]["d_hmax"]))
            dl_adj = float(sum(np.sum(gn[b] * dn[b]) for b in range(B)))
            checks.append({"tlm": dl_tlm, "adjoint": dl_adj, "rel_err": abs(dl_adj - dl_tlm) / abs(dl_tlm)})
        log(f"bench: field ensemble TLM checks {[round(c['rel_err'], 16) for c in checks]}")
        out.update({"forward_s": t_fwd, "gradient_s": t_grad, "gradient_first_call_s": gt[0],
                    "adjoint_forward_ratio": t_grad / t_fwd, "grads_per_s": B / t_grad,
                    "loss": loss, "tlm_checks": checks,
                    "speedup_vs_fd_per_cell": 2.0 * mem[0].mesh.nelt * t_fwd / B / (t_grad / B)})
    return out


def bench_scenarios(args):
    """64 calibration scenarios on one 100k-cell reach in one batch: cost of
    the batched zonal gradient (through the autograd Function) against the
    batched forward and against central differences with plain forwards;
    the gradient against central differences (recorded dt) and
    tangent-linear directional derivatives."""
    import numpy as np
    import torch

    from cuteflow_amd import calibration as CAL
    from cuteflow_amd import scenarios as SC
    from cuteflow_amd.adjoint import AdjointRun, GaugeLevels
    from cuteflow_amd.solver import BatchSolver, Solver

    nsteps, every, K, nz = 2000, 10, 100, 8
    sc = SC.calibration_batch(args.nscen, 100_000, ngauges=20, iflux=1)
    B = len(sc)
    meshes, params, U0 = [x.mesh for x in sc], [x.params for x in sc], [x.state for x in sc]
    n0 = [1.2 * x.n_zones for x in sc]
    mann0 = [n0[b][sc[b].zone] for b in range(B)]
    out = {"workload": f"{B} calibration scenarios (flood events) on one 100k-cell reach, water in the channel and "
                       "the floodplain dry at the start: own peak inflow 500-4000 m3/s, own hidden n in 8 zones "
                       "0.025-0.05, 20 channel gauges each, levels every 10 steps over 2000 steps, checkpoints "
                       "every 100 steps, HLLC + semi-implicit friction, zonal gradient at +20 % through the "
                       "autograd Function",
           "scenarios": B, "cells": sum(m.nelt for m in meshes), "scheme": "iflux 1 (HLLC), fricimplic 1"}
    for b in range(B):
        meshes[b].manning = mann0[b]
    options, out["schedule"] = section_options(args, meshes, params, U0, "scenarios")
    dev = torch.device("cuda")
    with BatchSolver(meshes, params, options=options) as bs:
        run = AdjointRun(bs, [x.gauges for x in sc], sample_every=every, checkpoint_every=K)
        obs, _ = run.forward(nsteps, U0=U0, manning=[x.n_zones[x.zone] for x in sc])
        for b in range(B):
            bs.set_manning(b, mann0[b])
        ts = []
        for _ in range(2):
            for b in range(B):
                bs.set_state(b, U0[b])
            sync()
            t = time.perf_counter()
            _, done, st = bs.step(nsteps)
            sync()
            ts.append(time.perf_counter() - t)
        assert int(done.min()) == nsteps and int(st.max()) == 0
        t_fwd = ts[-1]
        U0d = [torch.as_tensor(u, device=dev) for u in U0]   # initial states resident on the GPU
        zidx = torch.tensor(np.concatenate([nz * b + sc[b].zone for b in range(B)]), device=dev)
        obs_t = torch.tensor(np.concatenate([o.reshape(-1) for o in obs]), dtype=torch.float64, device=dev)

        def gradient():
            nzt = torch.tensor(np.concatenate(n0), dtype=torch.float64, device=dev, requires_grad=True)
            eta = GaugeLevels.apply(nzt[zidx], run, nsteps, U0d)
            loss = ((eta - obs_t) ** 2).sum()
            loss.backward()
            return nzt.grad.view(B, nz).cpu().numpy(), float(loss)

        gt = []
        for _ in range(2):
            sync()
            t = time.perf_counter()
            g, loss = gradient()
            sync()
            gt.append(time.perf_counter() - t)
        t_grad = gt[-1]
        log(f"bench: scenarios batched forward {t_fwd*1e3:.0f} ms, gradient {t_grad*1e3:.0f} ms, "
            f"ratio {t_grad / t_fwd:.2f}")
        # central differences with the plain forward: 2 x 8 runs per scenario
        with Solver(sc[0].mesh, sc[0].params) as s1:
            s1.set_manning(mann0[0])
            t1 = []
            for _ in range(3):
                s1.set_state(U0[0])
                sync()
                t = time.perf_counter()
                s1.step(nsteps)
                sync()
                t1.append(time.perf_counter() - t)
        t_plain = min(t1[1:])
        t_fd = 2 * nz * B * t_plain
        log(f"bench: scenarios plain single-scenario forward {t_plain*1e3:.1f} ms; central differences "
            f"{2 * nz * B} runs = {t_fd:.1f} s, {t_fd / t_grad:.1f}x the batched gradient")
        # checks: central differences at the recorded dt (all scenarios at
        # once, one zone and sign per tangent-linear value run) and
        # tangent-linear derivatives along a random zonal direction
        checks = {}
        for eps in (1e-6, 1e-7):
            vals = np.zeros((2, B, nz))
            # branch masks of the run at n0, then per (scenario, zone) the
            # (step, cell) updates that the two perturbed runs take
            # differently (wet/dry clamp; any branch)
            run.set_manning(mann0)
            run.set_mask_mode(3)
            run.tlm()
            run.set_mask_mode(4)
            run.flips(0, reset=True)
            run.flips(1, reset=True)
            cflips = np.zeros((B, nz), np.int64)
            bflips = np.zeros((B, nz), np.int64)
            for z in range(nz):
                for si, sgn in enumerate((1.0, -1.0)):
                    nzp = [x.copy() for x in n0]
                    for b in range(B):
                        nzp[b][z] *= 1.0 + sgn * eps
                    run.set_manning([nzp[b][sc[b].zone] for b in range(B)])
                    tl = run.tlm()
                    vals[si, :, z] = [CAL.gauge_misfit(tl[b]["eta"].reshape(-1), obs[b].reshape(-1))
                                      for b in range(B)]
                cflips[:, z] = [int(f.sum()) for f in run.flips(0, reset=True)]
                bflips[:, z] = [int(f.sum()) for f in run.flips(1, reset=True)]
            run.set_mask_mode(0)
            run.set_manning(mann0)
            gfd = (vals[0] - vals[1]) / (2.0 * eps * np.array(n0))
            scale = np.max(np.abs(gfd), axis=1, keepdims=True)
            err = np.abs(g - gfd) / np.maximum(scale, 1e-300)
            rel = np.abs(g - gfd) / np.maximum(np.abs(gfd), 1e-300)
            free = bflips == 0
            checks[f"fd_{eps:g}"] = {"max_rel_err_normwise": float(err.max()),
                                     "median_rel_err_normwise": float(np.median(err)),
                                     "fraction_within_1e-6": float(np.mean(err <= 1e-6)),
                                     "gradients_with_clamp_flips": int(np.sum(cflips > 0)),
                                     "gradients_with_branch_flips": int(np.sum(bflips > 0)),
                                     "branch_flip_free_gradients": int(np.sum(free)),
                                     "max_rel_err_branch_flip_free": float(rel[free].max()) if free.any() else None,
                                     "misses_without_branch_flips": int(np.sum((rel > 1e-6) & free))}
        rng = np.random.default_rng(5)
        dz = [rng.normal(size=nz) * x for x in n0]
        tl = run.tlm([dz[b][sc[b].zone] for b in range(B)])
        d_tlm = np.array([np.sum(2.0 * (tl[b]["eta"] - obs[b]) * tl[b]["d_eta"]) for b in range(B)])
        d_adj = np.array([np.dot(g[b], dz[b]) for b in range(B)])
        checks["tlm_max_rel_err"] = float(np.max(np.abs(d_adj - d_tlm) / np.abs(d_tlm)))
        log(f"bench: scenarios checks {checks}")
        out.update({"batched_forward_s": t_fwd, "gradient_s": t_grad, "gradient_first_call_s": gt[0],
                    "adjoint_forward_ratio": t_grad / t_fwd, "grads_per_s": B / t_grad,
                    "plain_single_forward_s": t_plain, "fd_plain_forward_s": t_fd,
                    "speedup_vs_fd_plain_forward": t_fd / t_grad,
                    "fd_cost_definition": f"central differences: 2 plain single-scenario forwards per zone and "
                                          f"scenario ({2 * nz * B} runs), each timed as one such run",
                    "loss": loss, "checks": checks})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--nx", type=int, default=3200)
    ap.add_argument("--ny", type=int, default=320)
    ap.add_argument("--order", choices=["rcm", "natural", "random"], default="rcm")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target duration of the CPU-baseline sample (0 disables)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--options", default="auto",
                    help="'auto' (the autotuner picks the schedule) or cf_params.options bits: 1 keep the input "
                         "cell order, 2 run every tile, 4 / 8 tiles of 128 / 64 cells, 16 XCD-contiguous tile order")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="before the timed region, run a device copy of this many MiB (3 times): a "
                         "known-size kernel that calibrates HBM byte counters in a profile")
    ap.add_argument("--tune-steps", type=int, default=100, help="steps per autotuner candidate")
    ap.add_argument("--extra-runs", type=int, default=1,
                    help="0: leave out the every-tile and random-order runs of the headline (for profiles of the "
                         "timed steps alone; needs --cpu-seconds 0)")
    ap.add_argument("--sections", default="c1,calibration,ensemble,field,scenarios",
                    help="comma-separated extra sections (c1, calibration, ensemble, field, scenarios); empty "
                         "for none")
    ap.add_argument("--nscen", type=int, default=64, help="scenarios of the scenarios section")
    ap.add_argument("--lbfgs-iters", type=int, default=20)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the time step runs on one subdomain per GPU and "
              "this build has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    torch.cuda.init()
    result = bench_headline(args)
    sections = [x for x in args.sections.split(",") if x]
    if "c1" in sections:
        result["c1"] = bench_c1(args)
    if "calibration" in sections:
        result["calibration"] = bench_calibration(args)
    if "ensemble" in sections:
        result["ensemble"] = bench_ensemble(args)
    if "scenarios" in sections:
        result["scenarios"] = bench_scenarios(args)
    if "field" in sections:
        result["batched_gradients"] = bench_field(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
