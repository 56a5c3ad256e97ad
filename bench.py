"""Benchmark: 2-D lid-driven cavity, 2049 x 2049 FP64, 50 Jacobi sweeps per step.

One step = one call of the cavity kernel (b, 1 + 50 pressure sweeps with four
boundary classes each, u, v, walls, lid) on the MI355X backend, fields resident
in HBM.  Prints one JSON line:

* value: cell-updates/s = 2049^2 x steps / wall time of the timed steps;
* roofline: the dominant kernel, the temporally blocked pressure loop (the
  Jacobi sweep and its boundary statements, up to 10 iterations per launch) -
  algorithmic bytes per launch (per iteration: class cells x 8 B x (arrays read
  + 1)) over its average launch duration, measured with HIP events on the
  backend stream over back-to-back launches, against 8 TB/s; `traffic` is the
  HBM bytes per launch from the committed rocprofv3 counter summary, with the
  measured HBM rate and the VALU figure of the committed SQ-counter summary
  beside it (the blocked kernel keeps its field on chip across iterations, so
  its algorithmic rate can exceed the HBM peak);
* cpu_baseline: the oracle's restatement of the reference CPU path (gcc
  -O3 -fopenmp) on a bounded sample of the same workload on the host cores.

``--dt 1e-5`` is the configured time step; it exceeds the explicit scheme's
stability limit, so fields overflow after a few hundred steps.  Timing does not
depend on the values; parity tests use dt=2e-6.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=2049)
    ap.add_argument("--dt", type=float, default=1e-5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU sample")
    ap.add_argument("--instrumented-steps", type=int, default=20)
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-fuse", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the device copy kernel on MIB MiB (byte-counter calibration)")
    ap.add_argument("--legs", default="all",
                    help="secondary workloads after the headline one: all, none, or a comma list")
    ap.add_argument("--batch-steps", type=int, default=100, help="steps of the batching leg")
    ap.add_argument("--heat-steps", type=int, default=1000, help="calls of the 3-D heat leg")
    ap.add_argument("--batch-no-plain", action="store_true",
                    help="batching leg: time the batched steps only (no plain comparison), e.g. under a profiler")
    ap.add_argument("--batch-max-blocks", type=int, default=0,
                    help="persistent grid cap of batched launches (0: one workgroup per tile)")
    ap.add_argument("--serve-requests", type=int, default=48, help="requests of the serving trace's prefix")
    ap.add_argument("--serve-full-requests", type=int, default=400,
                    help="requests of the whole serving trace, served after the prefix (0: skip)")
    ap.add_argument("--tune-limit", type=int, default=24,
                    help="instances the autotuner times per kernel in the batching/serving legs (0: no tuning)")
    ap.add_argument("--serve-budget-gib", type=float, default=16.0)
    ap.add_argument("--serve-tick", type=int, default=8, help="kernel calls per serving tick")
    ap.add_argument("--serve-max-mcells", type=float, default=64.0,
                    help="serving: cap on the admitted requests' cells, in millions (0: the HBM budget alone)")
    ap.add_argument("--serve-low-rate", type=float, default=2.0, help="arrival rate of the below-saturation run")
    ap.add_argument("--serve-low-requests", type=int, default=12)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; multi-GPU runs are not part of this build"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import numpy as np
    import xgrid_amd
    from xgrid_amd import _lib
    from xgrid_amd import workloads as W

    xgrid_amd.init(precision="double", cacheroot=os.path.join(ROOT, ".xgrid"),
                   graphs=not args.no_graphs, fuse=not args.no_fuse)
    n = args.n
    k = W.make_cavity()
    cfg = W.cavity_config(n, dt=args.dt)
    masks = W.cavity_masks(n)
    grids = [xgrid_amd.Grid((n, n), float) for _ in "bpuv"]
    for g, c in zip(grids, "bpuv"):
        g.boundary[...] = masks[c]
    lib = _lib.load()

    def step():
        k(*grids, cfg)

    for _ in range(max(1, args.warmup)):
        step()
    native = k.native
    _lib.check(lib.xg_synchronize())
    native.statement_counts(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    _lib.check(lib.xg_synchronize())
    t1 = time.perf_counter()
    wall = t1 - t0
    counts = native.statement_counts(reset=True)
    ms_per_step = wall / args.steps * 1e3
    cells = n * n
    value = cells * args.steps / wall
    step_bytes = native.algorithmic_bytes(counts, grids) / args.steps

    # dominant kernel: the Jacobi sweep (the implicit statement inside the loop)
    jac = [m.stmt.sid for m in native.lowered.stmts if m.stmt.mode == "pingpong"][-1]
    jm = native.lowered.stmts[jac]
    p_grid = grids[native.grid_index[jm.stmt.grid]]
    jac_cells = p_grid.class_info(jm.stmt.klass)["count"]
    jac_bytes = jac_cells * 8 * (len([r for r in jm.refs if r[1] >= 0]) + 1)
    # (a) events around every launch of the sweep during whole steps (graphs off)
    native.set_options(graphs=False)
    native.time_statement(jac, True)
    for _ in range(args.instrumented_steps):
        step()
    total_ms, launches = native.statement_time(jac)
    native.time_statement(jac, False)
    native.set_options()
    ev_ms = total_ms / max(1, launches)
    # (b) the sweep launched back to back between two events: kernel time alone
    reps = 200
    sweep_ms = native.time_isolated(jac, reps)
    avg_ms, bytes_per_launch = sweep_ms, jac_bytes
    kernel_name = f"xg_k{jac}_pp (Jacobi sweep)"
    kernel_id = f"xg_k{jac}_pp:n{n}"
    loop = native.loop_info(0) if native.lowered.loops else None
    if loop and loop["state"] == 1:
        # the pressure loop runs temporally blocked: its kernel dominates the step
        L = native.lowered.loops[0]
        per_iter = 0
        for m in L.stmts:
            g = grids[native.grid_index[m.stmt.grid]]
            per_iter += g.class_info(m.stmt.klass)["count"] * 8 * (len([r for r in m.refs if r[1] >= 0]) + 1)
        trip = 50                                   # the pressure loop's iterations per call
        launches_per_trip = -(-trip // L.tmax)
        avg_ms = native.time_loop(0, trip, reps=20)
        bytes_per_launch = per_iter * trip / launches_per_trip
        kernel_name = (f"{L.kernel} (pressure loop blocked: {trip} iterations of Jacobi sweep + 4 boundary "
                       f"statements in {launches_per_trip} launches of <= {L.tmax}, {L.core[0]}x{L.core[1]} "
                       f"tiles, halo {L.halo})")
        kernel_id = f"{L.kernel}:{L.core[0]}x{L.core[1]}:h{L.halo}:t{L.tmax}:w{L.threads}:n{n}"

    name = ctypes_name(lib)
    out = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"lid-driven cavity {n}x{n} FP64, 1+50 Jacobi sweeps/step, dt={args.dt}, "
                               "plain @xgrid.kernel calls",
                   "grid": [n, n], "sweeps_per_step": 51, "device": name},
        "roofline": roofline_entry(kernel_id, kernel_name, bytes_per_launch, avg_ms),
        "jacobi_sweep": {"kernel": f"xg_k{jac}_pp", "bytes_per_launch": jac_bytes, "avg_launch_ms": sweep_ms,
                         "GBs": jac_bytes / (sweep_ms * 1e-3) / 1e9,
                         "avg_launch_ms_in_steps": ev_ms, "launches_in_steps": launches},
        "blocked_loop": loop,
        "step_roofline": {"algorithmic_bytes_per_step": step_bytes,
                          "achieved_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9,
                          "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "counters": native.counters(),
    }
    if args.legs != "none":
        legs = {}
        wanted = set(LEGS) if args.legs == "all" else set(args.legs.split(","))
        for name, fn in LEGS.items():
            if name in wanted:
                print(f"# leg {name}", file=sys.stderr, flush=True)
                legs[name] = fn(args)
        out["legs"] = legs
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.dt, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def _timed(step, steps: int, warmup: int) -> float:
    """Seconds for `steps` calls of `step` after `warmup` untimed ones, device-synchronised."""
    from xgrid_amd import _lib
    lib = _lib.load()
    for _ in range(warmup):
        step()
    _lib.check(lib.xg_synchronize())
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    _lib.check(lib.xg_synchronize())
    return time.perf_counter() - t0


def leg_diffusion_1d(args) -> dict:
    """1-D diffusion, 4096 cells FP64, 500 calls: launch-bound (64 KiB of data)."""
    import xgrid_amd
    from xgrid_amd import workloads as W
    xgrid_amd.init(precision="double", cacheroot=os.path.join(ROOT, ".xgrid"))
    k = W.make_diffusion_1d()
    u0, mask, dx = W.diffusion_1d_setup(4096)
    nu = 0.3
    dt = 0.2 * dx * dx / nu
    u = xgrid_amd.Grid((4096,), float)
    u.now[...] = u0
    u.boundary[...] = mask
    steps = 500
    wall = _timed(lambda: k(u, nu, dt, dx), steps, 10)
    return {"workload": "1-D diffusion nx=4096 FP64, 500 calls", "us_per_call": wall / steps * 1e6,
            "cell_updates_per_s": 4096 * steps / wall, "counters": k.native.counters()}


def leg_cavity_fp32(args) -> dict:
    """The 2049^2 cavity with FP32 fields: throughput at the configured dt, and the
    FP32-vs-FP64 error and FP64 Poisson residuals at a stable dt."""
    import xgrid_amd
    from xgrid_amd import workloads as W
    from xgrid_amd.precision import CavityTwin
    n = args.n
    xgrid_amd.init(precision="float", cacheroot=os.path.join(ROOT, ".xgrid"))
    k = W.make_cavity()
    cfg = W.cavity_config(n, dt=args.dt)
    masks = W.cavity_masks(n)
    grids = [xgrid_amd.Grid((n, n), float) for _ in "bpuv"]
    for g, c in zip(grids, "bpuv"):
        g.boundary[...] = masks[c]
    steps = min(args.steps, 300)
    wall = _timed(lambda: k(*grids, cfg), steps, 5)
    step_bytes = 1312 // 2 * (n - 2) ** 2            # FP64's 1312 B per interior cell, halved
    out = {"workload": f"lid-driven cavity {n}x{n} FP32 fields, dt={args.dt}", "steps": steps,
           "ms_per_step": wall / steps * 1e3, "cell_updates_per_s": n * n * steps / wall,
           "step_roofline_frac": step_bytes / (wall / steps) / 1e9 / HBM_PEAK_GBS}
    del grids
    twin_steps = 50
    twin = CavityTwin(n, dt=2e-6)
    twin.run(twin_steps)
    out["precision"] = dict(twin.report(), dt=2e-6,
                            note="error = FP32 field - FP64 field over the interior, FP64 norms on the device; "
                                 "residual = FP64 norm of lap(p) - b of each run")
    return out


def leg_heat_3d(args) -> dict:
    """3-D heat, 512^3 FP32 storage, 1000 calls, then L2/Linf against the FP64 twin."""
    from xgrid_amd.precision import Heat3DTwin
    n, steps = 512, args.heat_steps
    twin = Heat3DTwin(n, seed=0)
    k32 = twin.kernels["float"]
    g32 = twin.fields["float"]
    warm = 3        # the first call uploads the 512^3 field and plans its mask classes on the host
    wall = _timed(lambda: k32(g32, twin.nu, twin.dt, twin.dx), steps, warm)
    t64 = _timed(lambda: twin.kernels["double"](twin.fields["double"], twin.nu, twin.dt, twin.dx), steps, warm)
    twin.steps = steps + warm
    sid = 0
    ms = k32.native.time_isolated(sid, 50)
    ms64 = twin.kernels["double"].native.time_isolated(sid, 50)
    kb = (n - 2) ** 3 * 4 * 2
    kb64 = (n - 2) ** 3 * 8 * 2
    return {"workload": f"3-D heat {n}^3 FP32 storage, 7-point FTCS, {steps} calls",
            "ms_per_step": wall / steps * 1e3, "cell_updates_per_s": n ** 3 * steps / wall,
            "fp64_ms_per_step": t64 / steps * 1e3,
            "roofline": {"kernel": f"xg_k{sid} (interior 7-point update)", "bytes_per_launch": kb,
                         "avg_launch_ms": ms, "achieved": kb / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "frac": kb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "fp64_roofline": {"kernel": f"xg_k{sid} of the FP64 twin (staged loads)", "bytes_per_launch": kb64,
                              "avg_launch_ms": ms64, "achieved": kb64 / (ms64 * 1e-3) / 1e9,
                              "peak": HBM_PEAK_GBS, "frac": kb64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "precision": twin.report()}


def leg_batching(args) -> dict:
    """256 independent cavity / 2-D diffusion instances (side log-uniform in [64, 1024],
    seed 1): K steps of all of them batched against the same K steps run instance by
    instance with plain kernel calls; results must be identical."""
    import numpy as np
    import xgrid_amd
    from xgrid_amd import workloads as W
    from xgrid_amd.multi import MultiSolver
    from xgrid_amd import _lib
    lib = _lib.load()
    specs = W.batch_specs(256, seed=1)
    K = args.batch_steps
    # Both sides are timed in steady state and take the same number of steps (the
    # results are compared).  A batch's time-level state repeats with a period of
    # up to 6 steps and each state's graph is captured on its second sighting: 13
    # warm-up steps reach replay.  Each plain instance is warmed with its 13 calls
    # in a row, which also keeps the plain path's graphs on (they switch off after
    # 64 calls without a single repeat).
    warm = 13
    cells = sum(s.n * s.n for s in specs)
    out = {"workload": f"{len(specs)} instances ({sum(s.kind == 'cavity' for s in specs)} cavity, "
                       f"{sum(s.kind == 'diffusion' for s in specs)} 2-D diffusion), side log-uniform [64, 1024], "
                       f"seed 1, {K} steps", "cells": cells}
    for prec in ("double", "float"):
        ms = MultiSolver(prec, max_blocks=args.batch_max_blocks)
        tuning = ms.autotune(specs, limit=args.tune_limit) if args.tune_limit else {}
        handles = [ms.add(s) for s in specs]
        ms.step(warm)
        _lib.check(lib.xg_synchronize())
        t0 = time.perf_counter()
        ms.step(K)
        _lib.check(lib.xg_synchronize())
        t_batch = time.perf_counter() - t0
        print(f"# batching {prec}: batched steps done", file=sys.stderr, flush=True)
        key = "f64" if prec == "double" else "f32"
        if args.batch_no_plain:
            out[key] = {"batched_s": t_batch, "ms_per_batched_step": t_batch / K * 1e3,
                        "cell_updates_per_s_batched": cells * K / t_batch, "counters": ms.counters(),
                        "tile_shapes": {k: list(v) if v else None for k, v in ms.shapes.items()}}
            ms.close()
            continue
        plain_before = {k: op.native.counters() for k, op in ms.ops.items()}
        # The plain side is timed twice, with graph replay on (the default) and off, and
        # the faster is the baseline: going from instance to instance, each instance's
        # graphs are captured again after its predecessors evicted them, which costs the
        # replays' gain when the device is the bottleneck.
        # With graph replay switched off process-wide (XGRID_GRAPHS=0), only the graphs-off
        # pass runs.
        plain_times = {}
        for graphs in ((True, False) if xgrid_amd.get_config().graphs else (False,)):
            for op in ms.ops.values():
                if op.native is not None:
                    op.native.set_options(graphs=graphs)
            plain = []
            for s in specs:
                grids, scalars = ms.grids_for(s)
                for _ in range(warm):
                    ms.plain_call(s, grids, scalars)
                plain.append((grids, scalars))
            _lib.check(lib.xg_synchronize())
            t0 = time.perf_counter()
            for s, (grids, scalars) in zip(specs, plain):
                for _ in range(K):
                    ms.plain_call(s, grids, scalars)
            _lib.check(lib.xg_synchronize())
            plain_times["graphs_on" if graphs else "graphs_off"] = time.perf_counter() - t0
            print(f"# batching {prec}: plain calls done (graphs {'on' if graphs else 'off'})",
                  file=sys.stderr, flush=True)
            same = all(np.array_equal(a.now, b.now) for h, (grids, _) in zip(handles, plain)
                       for a, b in zip(ms.grids(h), grids))
            if not same:
                break
            del plain
        for op in ms.ops.values():
            if op.native is not None:
                op.native.set_options()
        t_plain = min(plain_times.values())
        out[key] = {
            "batched_s": t_batch, "plain_s": t_plain, "plain_s_by_graphs": plain_times, "speedup": t_plain / t_batch,
            "solves_per_s_batched": len(specs) / t_batch, "solves_per_s_plain": len(specs) / t_plain,
            "cell_updates_per_s_batched": cells * K / t_batch, "ms_per_batched_step": t_batch / K * 1e3,
            "identical_to_plain": bool(same), "counters": ms.counters(),
            "tile_shapes": {k: list(v) if v else None for k, v in ms.shapes.items()},
            "autotune": {k: {"timings_ms_per_step": v["timings"], "cached": v["cached"]} for k, v in tuning.items()},
            "plain_counters": {k: {c: v - plain_before[k][c] for c, v in op.native.counters().items()}
                               for k, op in ms.ops.items()}}
        ms.close()
    out["note"] = ("a solve = K steps of one instance; plain = each instance's kernel called K times, one instance "
                   "after another, the faster of graph replay on and off")
    return out


def leg_serving(args) -> dict:
    """Poisson arrivals (200/s, seed 2) of FP32 cavity solves, side log-uniform in
    [128, 2048], 100-2000 steps, served with continuous batching under an HBM budget;
    the same trace served one request at a time with plain calls; and a lower
    arrival rate below saturation."""
    import numpy as np
    from xgrid_amd import workloads as W
    from xgrid_amd.serving import Server, serve_plain
    budget = int(args.serve_budget_gib * (1 << 30))
    trace = W.serving_specs(args.serve_requests, seed=2, rate=200.0)
    srv = Server("float", hbm_budget=budget, steps_per_tick=args.serve_tick,
                 max_cells=int(args.serve_max_mcells * 1e6) if args.serve_max_mcells else None)
    tuning = srv.ms.autotune(trace, limit=args.tune_limit) if args.tune_limit else {}
    def progress(label, total):
        def on_done(_, n):
            if n % 25 == 0 or n == total:
                print(f"# serving {label}: {n} of {total} requests done", file=sys.stderr, flush=True)
        return on_done

    def served_and_plain(tr, label):
        got, summ = srv.run(tr, on_done=progress(f"batched{label}", len(tr)))
        summ.update(peak_resident_requests=srv.peak_active, peak_resident_bytes=srv.peak_used, ticks=srv.ticks)
        ref, summ_plain = serve_plain(tr, "float", solver=srv.ms, on_done=progress(f"plain FIFO{label}", len(tr)))
        same = all(np.array_equal(got[r].fields[c], ref[r].fields[c]) for r in got for c in "bpuv")
        return summ, summ_plain, same

    summ, summ_plain, same = served_and_plain(trace, "")
    full = None
    if args.serve_full_requests > len(trace):
        # the whole trace the prefix is taken from, as a second entry of the leg
        tr_full = W.serving_specs(args.serve_full_requests, seed=2, rate=200.0)
        f_summ, f_plain, f_same = served_and_plain(tr_full, " (full trace)")
        full = {"requests": len(tr_full), "cell_steps": sum(s.n * s.n * s.steps for s in tr_full),
                "batched": f_summ, "plain_fifo": f_plain,
                "speedup_solves_per_s": f_summ["solves_per_s"] / f_plain["solves_per_s"],
                "identical_to_plain": bool(f_same)}
    low = W.serving_specs(args.serve_low_requests, seed=3, rate=args.serve_low_rate)
    got_low, summ_low = srv.run(low)
    summ_low.update(peak_resident_requests=srv.peak_active)
    del got_low
    srv.close()
    work = sum(s.n * s.n * s.steps for s in trace)
    return {"workload": f"{len(trace)} requests of the seed-2 trace (Poisson 200/s, cavity FP32, side log-uniform "
                        f"[128, 2048], 100-2000 steps, nu in [0.001, 0.1]), HBM budget {args.serve_budget_gib} GiB, "
                        f"resident-cell cap {args.serve_max_mcells:g}M, {args.serve_tick} steps per tick",
            "cell_steps": work, "batched": summ, "plain_fifo": summ_plain,
            "speedup_solves_per_s": summ["solves_per_s"] / summ_plain["solves_per_s"],
            "identical_to_plain": bool(same), "tile_shapes": {k: list(v) if v else None for k, v in srv.ms.shapes.items()},
            "autotune": {k: {"timings_ms_per_step": v["timings"], "cached": v["cached"]} for k, v in tuning.items()},
            "below_saturation": dict(summ_low, rate_per_s=args.serve_low_rate,
                                     workload=f"{len(low)} requests, seed 3, Poisson {args.serve_low_rate}/s"),
            "full_trace": full}


def leg_compile(args) -> dict:
    """First-call latency of a kernel the process has not built yet (the 2049^2 FP64
    cavity under a name and host macro of its own, so that neither the on-disk cache
    nor the compiler's cache holds it): cold (hipRTC and the host stub compile inside
    the first call), warm (the same text again: both come from the on-disk cache) and
    hidden (the compile starts in the background when the kernel is decorated, while
    the caller allocates and fills its grids, as a script does)."""
    import uuid
    import xgrid_amd
    from xgrid_amd import _lib
    from xgrid_amd import workloads as W
    from xgrid_amd.operator import Operator
    xgrid_amd.init(precision="double", cacheroot=os.path.join(ROOT, ".xgrid"), precompile=False)
    lib = _lib.load()
    n = args.n
    func = W.make_cavity().func
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)

    def setup():
        grids = [xgrid_amd.Grid((n, n), float) for _ in "bpuv"]
        for g, c in zip(grids, "bpuv"):
            g.boundary[...] = masks[c]
        return grids

    def first_call(op, grids) -> float:
        t0 = time.perf_counter()
        op(*grids, cfg)
        _lib.check(lib.xg_synchronize())
        return time.perf_counter() - t0

    tag = uuid.uuid4().hex[:12]
    grids = setup()
    cold = first_call(Operator(func, "kernel", f"cavity_cold_{tag}", macro=[f"/* {tag} */"]), grids)
    warm = first_call(Operator(func, "kernel", f"cavity_cold_{tag}", macro=[f"/* {tag} */"]), grids)
    del grids
    t0 = time.perf_counter()
    op = Operator(func, "kernel", f"cavity_hidden_{tag}", macro=[f"/* {tag} hidden */"])
    pre = op.precompile()                         # what @kernel() does when the configuration is set
    grids = setup()
    setup_s = time.perf_counter() - t0
    hidden = first_call(op, grids)
    # decorated long before its first call (at import, say): the compile is over by then
    op2 = Operator(func, "kernel", f"cavity_early_{tag}", macro=[f"/* {tag} early */"])
    pre2 = op2.precompile(wait=True)
    early = first_call(op2, grids)
    return {"workload": f"first call of a new {n}x{n} FP64 cavity kernel (unique source text)",
            "cold_first_call_s": cold, "warm_first_call_s": warm,
            "hidden_during_setup": {"first_call_s": hidden, "setup_s_before_call": setup_s,
                                    "background_compile_s": pre.seconds,
                                    "compile_error": repr(pre.error) if pre.error else None},
            "compiled_before_call": {"first_call_s": early, "background_compile_s": pre2.seconds},
            "note": "each figure includes the call's own device work (one cavity step, plus the blocked loop's "
                    "validation launch) and the module load"}


LEGS = {"diffusion_1d": leg_diffusion_1d, "cavity_fp32": leg_cavity_fp32, "heat_3d_fp32": leg_heat_3d,
        "batching": leg_batching, "serving": leg_serving, "compile": leg_compile}


def copy_probe(mib: int, iters: int = 20) -> None:
    """Copy `mib` MiB on the device `iters` times (read + write of every byte)."""
    import ctypes
    from xgrid_amd import _lib
    lib = _lib.ensure_device(0)
    g, ms = ctypes.c_double(), ctypes.c_double()
    _lib.check(lib.xg_copy_probe(mib << 20, iters, ctypes.byref(g), ctypes.byref(ms)))
    print(json.dumps({"copy_probe_MiB": mib, "iters": iters, "GBs": g.value, "ms_per_copy": ms.value,
                      "bytes_read_per_copy": mib << 20, "bytes_written_per_copy": mib << 20}), flush=True)


def roofline_entry(kernel_id: str, kernel_name: str, bytes_per_launch: float, avg_ms: float) -> dict:
    """The dominant kernel against the HBM roofline.  `achieved` is the algorithmic
    (unblocked-sweep) bytes per launch over the launch time; a temporally blocked
    kernel can exceed the peak on that count, so the measured HBM bytes, their rate
    and the VALU figure from the committed counter summaries stand beside it."""
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic(kernel_id)
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel_id": kernel_id, "kernel": kernel_name,
           "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms}
    if traffic:
        out["hbm_measured_GBs"] = traffic / (avg_ms * 1e-3) / 1e9
        out["hbm_measured_frac"] = out["hbm_measured_GBs"] / HBM_PEAK_GBS
    sq = measured_counters(kernel_id)
    if sq:
        out["valu_busy"] = sq.get("valu_busy")
        out["lds_wait_frac"] = sq.get("lds_wait_frac")
    if traffic and out["frac"] > 1.0:
        out["note"] = ("temporally blocked: 'achieved' counts every sweep's algorithmic bytes, while HBM moves "
                       "'traffic' bytes per launch; the kernel is bound on chip (VALU, LDS), see valu_busy")
    return out


def measured_counters(kernel: str):
    """SQ-counter figures of `kernel` from the committed rocprofv3 summary, or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "sq_counters.json")) as f:
            return json.load(f).get("kernels", {}).get(kernel)
    except (OSError, ValueError):
        return None


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    entry = data.get("kernels", {}).get(kernel)
    return None if entry is None else entry.get("bytes_per_launch")


def ctypes_name(lib) -> str:
    import ctypes
    buf = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    hbm = ctypes.c_int64()
    lib.xg_device_name(buf, 256, ctypes.byref(cus), ctypes.byref(hbm))
    return f"{buf.value.decode()} {cus.value} CUs {hbm.value / 2**30:.0f} GiB"


def cpu_baseline(n: int, dt: float, seconds: float) -> dict:
    """The oracle's restated reference CPU path on the same workloads, bounded
    samples: the FP64 cavity (the headline), the 1-D diffusion, the 512^3 FP32
    heat equation and the FP64 batching mix."""
    from xgrid_amd.precision import precision_scope
    with precision_scope("double"):        # the secondary legs may leave FP32 in force
        return _cpu_baseline(n, dt, seconds)


def _cpu_baseline(n: int, dt: float, seconds: float) -> dict:
    import xgrid_amd
    from xgrid_amd import workloads as W
    from oracle.cpu_path import CpuKernel, HostGrid
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    k = W.make_cavity()
    ck = CpuKernel(k, opt_level=3)
    ck.threads(threads)
    cfg = W.cavity_config(n, dt=dt)
    masks = W.cavity_masks(n)
    grids = [HostGrid((n, n), float) for _ in "bpuv"]
    for g, c in zip(grids, "bpuv"):
        g.boundary[...] = masks[c]
    ck(*grids, cfg)        # first call: JIT compile and first-touch of the levels
    steps = 0
    t0 = time.perf_counter()
    while True:
        ck(*grids, cfg)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    # the same path on one thread, a shorter sample
    ck.threads(1)
    steps1 = 0
    t0 = time.perf_counter()
    while True:
        ck(*grids, cfg)
        steps1 += 1
        el1 = time.perf_counter() - t0
        if el1 >= seconds / 3 or steps1 >= 10:
            break
    # the 1-D diffusion (4096 cells FP64) per call on the same path: Python, argument
    # serialisation and one OpenMP nest per statement, as the reference runs it
    kd = W.make_diffusion_1d()
    ckd = CpuKernel(kd, opt_level=3)
    ckd.threads(threads)
    u0, dmask, dx = W.diffusion_1d_setup(4096)
    ud = HostGrid((4096,), float)
    ud.now[...] = u0
    ud.boundary[...] = dmask
    ddt = 0.2 * dx * dx / 0.3
    for _ in range(20):
        ckd(ud, 0.3, ddt, dx)
    t0 = time.perf_counter()
    for _ in range(500):
        ckd(ud, 0.3, ddt, dx)
    d_us = (time.perf_counter() - t0) / 500 * 1e6
    del grids
    return {"value": n * n * steps / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the {n}x{n} cavity ({el:.1f} s), gcc -O3 -fopenmp, "
                      f"OMP_NUM_THREADS={threads}",
            "single_thread": {"value": n * n * steps1 / el1, "cores": 1,
                              "sample": f"{steps1} steps of the {n}x{n} cavity ({el1:.1f} s), one thread"},
            "diffusion_1d": {"us_per_call": d_us, "cores": threads,
                             "sample": "500 calls of the 1-D diffusion, 4096 cells FP64"},
            "heat_3d": cpu_heat_3d(threads, seconds / 3),
            "batching": cpu_batching(threads, seconds / 2)}


def cpu_heat_3d(threads: int, seconds: float, n: int = 512) -> dict:
    """The restated reference CPU path (true 3-D strides) on the 512^3 FP32 heat
    equation: a prefix of the 1000 steps, timed per step."""
    import numpy as np
    import xgrid_amd
    from xgrid_amd import workloads as W
    from xgrid_amd.precision import precision_scope
    from oracle.cpu_path import CpuKernel, HostGrid
    with precision_scope("float"):
        ck = CpuKernel(W.make_heat_3d(), opt_level=3)
        ck.threads(threads)
        u0, mask, dx = W.heat_3d_setup(n, np.float32, seed=0)
        nu = 0.1
        dt = 0.1 * dx * dx / nu
        u = HostGrid((n, n, n), float)
        u.now[...] = u0
        u.boundary[...] = mask
        del u0
        ck(u, nu, dt, dx)                     # JIT compile and first touch of the levels
        steps = 0
        t0 = time.perf_counter()
        while True:
            ck(u, nu, dt, dx)
            steps += 1
            el = time.perf_counter() - t0
            if el >= seconds or steps >= 20:
                break
    return {"s_per_step": el / steps, "cell_updates_per_s": n ** 3 * steps / el, "cores": threads, "kind": "port",
            "sample": f"{steps} of the 1000 steps of the {n}^3 FP32 heat equation ({el:.1f} s), gcc -O3 -fopenmp; "
                      f"the whole run extrapolates to {el / steps * 1000:.0f} s"}


def cpu_batching(threads: int, seconds: float) -> dict:
    """CPU solves/s of the batching mix (256 cavity / 2-D diffusion instances, seed
    1, FP64, a solve = 100 steps of one instance) on the restated reference CPU
    path: every 8th instance (sorted by kind, then size) for a few steps each,
    scaled to all 256 instances and 100 steps."""
    import xgrid_amd
    from xgrid_amd import workloads as W
    from oracle.cpu_path import CpuKernel, HostGrid
    specs = W.batch_specs(256, seed=1)
    kernels = {"cavity": CpuKernel(W.make_cavity(), opt_level=3), "diffusion": CpuKernel(W.make_diffusion_2d(), opt_level=3)}
    for ck in kernels.values():
        ck.threads(threads)
    order = sorted(specs, key=lambda s: (s.kind, s.n))
    sample = order[::8]
    # per-instance seconds per step, first call (first touch) excluded
    per_step = {}
    budget = seconds / len(sample)
    for s in sample:
        grids, scalars = W.instance_grids(s, HostGrid)
        ck = kernels[s.kind]
        ck(*grids, *scalars)
        k = 0
        t0 = time.perf_counter()
        while True:
            ck(*grids, *scalars)
            k += 1
            el = time.perf_counter() - t0
            if el >= budget or k >= 10:
                break
        per_step[s.rid] = el / k
    # each sampled instance stands for the 8 around it in the sorted order (same kind, near size)
    step_s = sum(per_step[s.rid] * len(order[i * 8:(i + 1) * 8]) for i, s in enumerate(sample))
    K = 100
    return {"solves_per_s": len(specs) / (step_s * K), "s_per_step_all_instances": step_s, "cores": threads,
            "kind": "port", "steps_per_solve": K,
            "sample": f"{len(sample)} of the {len(specs)} instances (every 8th by kind and size), up to 10 steps "
                      f"each; each stands for the 8 instances around it; FP64, gcc -O3 -fopenmp"}


if __name__ == "__main__":
    main()
