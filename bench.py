#!/usr/bin/env python3
"""Benchmark of the FP64 finite-volume St-Venant time step on one MI355X.

Workload (the river-reach configuration): a 20 km x 2 km reach meshed as
3200 x 320 jittered quads split into 2,048,000 triangles in RCM order, bed
slope 1/2000 + 0.5 m sinusoidal bars (300 m) + a seeded Gaussian-smoothed
random field (sigma 0.3 m, 200 m correlation), a 3 m deep main channel in a
floodplain that starts dry, Manning n 0.03 / 0.06, inflow ramped 0 ->
3000 m3/s over 1 h, fixed outlet level, dry threshold 1e-3 m, WAF second
order with the superbee limiter and semi-implicit friction (fricimplic 1).

A step is one time step over the whole mesh (one launch of the fused step
kernel).  Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_WAF = 156  # algorithmic HBM bytes per cell-update
BYTES_HLLC = 132
PEAK_HBM_GBS = 8000.0
METRIC = "FP64 cell-updates/s + %HBM roofline; adjoint/forward cost ratio; batched grads/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(nx, ny, order, seed=1):
    import numpy as np

    from cuteflow_amd import mesh as M
    from cuteflow_amd.params import Params

    r = M.river_mesh(nx=nx, ny=ny, length=20000.0 * nx / 3200, width=2000.0 * ny / 320, seed=seed)
    m, U0 = r.mesh, r.state
    if order == "rcm":
        perm = M.rcm_order(m)
    elif order == "random":
        perm = M.random_order(m, seed)
    else:
        perm = None
    if perm is not None:
        m, U0 = m.permuted(perm), U0[perm]
    p = Params(iflux=3, ilimiteur=1, tolisec=1e-3, cfl=0.9, ts=1e300, inlet="inflow",
               outlet="fixedheight", friction=1, fricimplic=1, nombre_entrees=1, nombre_sorties=1)
    p.h_sortie[0] = r.h_out
    p.hydrographs = {1: ([0.0, r.t_ramp], [0.0, r.q_peak])}
    return m, p, np.ascontiguousarray(U0)


def traffic_from_profile(workload_key):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_key}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


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
    ap.add_argument("--options", type=int, default=0,
                    help="cf_params.options bits: 1 keep the input cell order, 2 run every tile")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="before the timed region, run a device copy of this many MiB (3 times): a "
                         "known-size kernel that calibrates HBM byte counters in a profile")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the time step runs on one subdomain per GPU and "
              "this build has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    torch.cuda.init()
    from cuteflow_amd import _lib
    from cuteflow_amd.solver import Solver

    t0 = time.perf_counter()
    m, p, U0 = make_workload(args.nx, args.ny, args.order)
    log(f"bench: workload nelt={m.nelt} order={args.order} built in {time.perf_counter() - t0:.1f}s")
    s = Solver(m, p, options=args.options)
    s.set_state(U0)
    lib = s.lib
    if args.warmup:
        s.step(args.warmup)
    log(f"bench: warmup {args.warmup} steps done, tc={s.tc:.3f}s")

    if args.copy_probe_mb:
        n = args.copy_probe_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: copy probe {args.copy_probe_mb} MiB read + {args.copy_probe_mb} MiB written, 3 times")
        del src, dst

    stream = torch.cuda.ExternalStream(s.stream)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    hist = np.zeros(args.steps + 1)
    import ctypes as C

    done = C.c_int()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    _lib.check(lib.cf_step_async(s.h, args.steps), "cf_step_async")
    ev1.record(stream)
    _lib.check(lib.cf_step_finish(s.h, _lib.dptr(hist), C.byref(done)), "cf_step_finish")
    torch.cuda.synchronize()
    w1 = time.perf_counter()
    if done.value != args.steps:
        raise RuntimeError(f"run stopped after {done.value} of {args.steps} steps")
    wall = w1 - w0
    kern_ms = ev0.elapsed_time(ev1)
    per_launch_s = kern_ms / 1000.0 / args.steps
    stats = s.last_run_stats()
    value = m.nelt * args.steps / wall
    # cells the step kernel actually processed per launch: dry tiles with dry
    # neighbours are skipped (their update is the identity)
    active_cells = stats["active_fraction"] * m.nelt
    achieved = BYTES_WAF * active_cells / per_launch_s / 1e9
    wet = int(np.sum(s.get_state()[:, 0] > p.tolisec))
    log(f"bench: {args.steps} steps in {wall:.3f}s wall, {kern_ms:.2f} ms on the stream, "
        f"{value:.3e} cell-updates/s, active tiles {stats['active_fraction']:.3f}, "
        f"{achieved:.0f} GB/s algorithmic over processed cells, wet cells {wet}")

    workload_key = f"river_waf_{m.nelt}_{args.order}"
    traffic = traffic_from_profile(workload_key)
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
            "device_order": "caller's" if args.options & 1 else "Hilbert curve of centroids",
            "dry_tile_skipping": not (args.options & 2),
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
            "kernel": "step_kernel<3,2>",
            "bytes_per_cell_update": BYTES_WAF,
            "cells_per_launch": active_cells,
            "active_tile_fraction": stats["active_fraction"],
            "avg_launch_us": per_launch_s * 1e6,
            "timing": "HIP events on the solver stream over the timed steps (step kernel + "
                      "one-workgroup finalize kernel per step)",
        },
    }

    # CPU baseline: the oracle on a bounded sample of the same workload, on
    # the host cores, started from the same initial state; the GPU result
    # after the same number of steps is compared bit for bit.
    if args.cpu_seconds > 0:
        from oracle import oracle as O

        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        t0 = time.perf_counter()
        prep = O.Prepared(m)
        log(f"bench: oracle preprocessing {time.perf_counter() - t0:.1f}s")
        t0 = time.perf_counter()
        O.run(prep, p, U0, 2, nthreads=threads)
        per_step = (time.perf_counter() - t0) / 2
        nsamp = max(2, min(args.steps, int(args.cpu_seconds / max(per_step, 1e-6))))
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
                      f"(C restatement of the reference kernels, OpenMP, {threads} threads, 1 MPI rank)",
        }
        result["parity_check"] = {"steps": nsamp, "bit_identical": same}
    s.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
