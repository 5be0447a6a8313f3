#!/usr/bin/env python3
"""Benchmark: full GARM-LS time steps of the 3D dam break on one MI355X.

Workload: 3D dam break, n^3 MAC grid (default 256^3 cells including a 2-cell
solid band; liquid box x < 0.4, y < 0.6, z < 0.5 of the unit domain), gravity
-9.8, CFL 1, frame rate 25 (the reference's dt rule), WENO5 reinitialisation
(30 iterations) every 5 steps, pressure PCG to relative residual 1e-6.
A step is one LevelSetLiquid::advance (LevelSetLiquid.cpp:608-698) with every
field resident in HBM.

Prints one JSON line: cell-updates/s over the timed steps, per-phase times,
the roofline position of the dominant kernel (timed live with HIP events on
the context's stream) and the CPU oracle timed on this host on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "garm-ls_amd")]

from garm_amd import capi, scenes  # noqa: E402
from garm_amd.driver import FrameClock  # noqa: E402

METRIC = "cell-updates/s at 256³, mixed+learned PCG vs plain FP64 PCG; HBM GB/s; vol err"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(args, mode_name):
    g = scenes.grid(args.n, 3)
    mode = {"fp64": capi.PCG_FP64, "mixed": capi.PCG_MIXED, "learned": capi.PCG_LEARNED}[mode_name]
    opts = capi.step_opts(reinit_interval=args.reinit_interval, tol=args.tol, max_iterations=args.max_iterations,
                          check_interval=args.check_interval, mode=mode)
    t0 = time.time()
    sim = capi.Liquid(3, g["res"], g["dx"], g["origin"], g["bw"])
    scenes.dam_break(sim, g)
    sim.initialize(opts)
    log(f"[bench] {mode_name}: {args.n}^3 initialised in {time.time() - t0:.1f}s")
    clk = FrameClock(g["dx"], args.cfl)
    step = 0
    for _ in range(args.warmup):
        sim.step(clk.next_dt(sim.max_speed()), step, opts)
        step += 1
    log(f"[bench] {args.warmup} warm-up steps done")
    sim.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        dt = clk.next_dt(sim.max_speed())
        stats.append(sim.step(dt, step, opts))
        step += 1
        if k % 5 == 4:
            log(f"[bench] timed step {k + 1}/{args.steps}")
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    return g, sim, stats, elapsed


def roofline(sim, stats, args):
    K = len(stats)
    iters = sum(s["pcg"]["iterations"] + s["pcg"]["outer_iterations"] for s in stats) / K    # SpMV launches per step
    mixed = args.mode != "fp64"
    reinit_stages = 3 * 30 * sum(s["reinitialized"] for s in stats) / K
    per_step = {capi.K_ADVECT_REFMAP: 1, capi.K_ADVECT_VELOCITY: 1, capi.K_COMPOSE: 1 + sum(s["restarted"] for s in stats) / K,
                capi.K_REINIT_STAGE: reinit_stages, capi.K_PCG_SPMV: 0 if mixed else iters, capi.K_PCG_UPDATE: 0 if mixed else iters,
                capi.K_PCG32_SPMV: iters if mixed else 0, capi.K_PCG32_UPDATE: iters if mixed else 0}
    kernels = {}
    for k in per_step:
        if per_step[k] == 0:
            continue
        ms, by = sim.time_kernel(k, reps=args.kernel_reps)
        kernels[capi.KERNEL_NAMES[k]] = dict(avg_ms=ms, bytes=by, gbs=by / ms / 1e6, per_step=per_step[k],
                                            ms_per_step=ms * per_step[k])
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    d = kernels[dom]
    return dom, kernels, {"kernel": dom, "bound": "hbm", "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                          "avg_launch_ms": round(d["avg_ms"], 5), "bytes_per_launch": d["bytes"]}


def cpu_baseline(args):
    """The oracle on this host: one full step of the same workload (no
    reinitialisation in it) plus the 30 / k WENO5 iterations a step carries
    on average at a reinit interval of k, after an un-timed set-up."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = args.cpu_n
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if args.cpu_threads > 0:
        threads = args.cpu_threads
    O.Oracle.set_threads(threads)
    g = scenes.grid(n, 3)
    o = O.Oracle(3, g["res"], g["dx"], g["origin"], g["bw"])
    scenes.dam_break(o, g)
    t0 = time.time()
    o.initialize(reinit_iters=0, precond=O.PRECOND_IC0, tol=args.tol, max_it=args.max_iterations)
    clk = FrameClock(g["dx"], args.cfl)
    dt0 = clk.next_dt(o.max_speed())
    o.step(dt0, do_reinit=False, precond=O.PRECOND_IC0, tol=args.tol, max_it=args.max_iterations)
    log(f"[bench] CPU oracle set-up {time.time() - t0:.1f}s")
    dt = clk.next_dt(o.max_speed())
    t0 = time.perf_counter()
    r = o.step(dt, do_reinit=False, precond=O.PRECOND_IC0, tol=args.tol, max_it=args.max_iterations)
    t_step = time.perf_counter() - t0
    reinit_iters = max(1, round(30 / args.reinit_interval))
    t0 = time.perf_counter()
    o.reinit(reinit_iters, 0.9)
    t_reinit = time.perf_counter() - t0
    total = t_step + t_reinit
    cells = n ** 3
    log(f"[bench] CPU oracle step {t_step:.1f}s + {reinit_iters} reinit iterations {t_reinit:.1f}s")
    return {"value": round(cells / total, 1), "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C++/OpenMP FP64 restatement, IC(0)-PCG tol {args.tol:g}): one {n}^3 dam-break step "
                      f"({r['iterations']} PCG iterations) + {reinit_iters} WENO5 iterations (30 every "
                      f"{args.reinit_interval} steps, amortised) = {total:.1f}s",
            "seconds": round(total, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--mode", choices=["fp64", "mixed", "learned"], default="mixed")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 run beside a mixed/learned one")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iterations", type=int, default=20000)
    ap.add_argument("--check-interval", type=int, default=16)
    ap.add_argument("--reinit-interval", type=int, default=5)
    ap.add_argument("--cfl", type=float, default=1.0)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU decomposition is designed "
              "but not built", file=sys.stderr)
        sys.exit(2)

    cells = args.n ** 3
    plain = None
    if args.mode != "fp64" and not args.no_plain:
        g, sim, stats, elapsed = run_gpu(args, "fp64")
        sim.close()
        K = len(stats)
        plain = {"value": round(cells * K / elapsed, 1), "ms_per_step": round(1000 * elapsed / K, 3),
                 "pcg_iterations_mean": sum(s["pcg"]["iterations"] for s in stats) / K,
                 "phases_ms": {k: round(sum(s[k] for s in stats) / K, 3) for k in ("t_advect_ms", "t_restart_ms", "t_project_ms", "t_extrap_ms")}}
    g, sim, stats, elapsed = run_gpu(args, args.mode)
    K = len(stats)
    value = cells * K / elapsed
    phases = {k: round(sum(s[k] for s in stats) / K, 3) for k in ("t_advect_ms", "t_restart_ms", "t_project_ms", "t_extrap_ms", "t_total_ms")}
    pcg_iters = [s["pcg"]["iterations"] for s in stats]
    dom, kernels, roof = roofline(sim, stats, args)
    sim.close()
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(1000 * elapsed / K, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.mode == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"3D dam break full step, {args.n}^3 MAC grid (liquid box x<0.4, y<0.6, z<0.5, gravity -9.8, CFL {args.cfl:g})",
                   "pcg": args.mode, "tolerance": args.tol, "reinit_interval": args.reinit_interval,
                   "cells": cells},
        "phases_ms": phases,
        "pcg_iterations": {"mean": sum(pcg_iters) / K, "min": min(pcg_iters), "max": max(pcg_iters),
                           "outer_mean": sum(s["pcg"]["outer_iterations"] for s in stats) / K},
        "pcg_residual_max": max(s["pcg"]["relative_residual"] for s in stats),
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(value / plain["value"], 3) if plain else None,
        "restarts": sum(s["restarted"] for s in stats),
        "kernels": {k: {kk: round(vv, 5) if isinstance(vv, float) else vv for kk, vv in v.items()} for k, v in kernels.items()},
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
