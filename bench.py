#!/usr/bin/env python3
"""Benchmark: full GARM-LS time steps of the 3D dam break on one MI355X.

Workload: 3D dam break, n^3 MAC grid (default 256^3 cells including a 2-cell
solid band; liquid box x < 0.4, y < 0.6, z < 0.5 of the unit domain), gravity
-9.8, CFL 1, frame rate 25 (the reference's dt rule), WENO5 reinitialisation
(30 iterations) every 5 steps, pressure solve to relative residual 1e-6.
--scene droplet runs the droplet-into-pool scene instead (pool y < 0.3,
sphere r = 0.08 at a seeded (x, 0.7, z), falling at 2).
A step is one LevelSetLiquid::advance (LevelSetLiquid.cpp:608-698) with every
field resident in HBM.

Configurations of the same workload, run one after the other:
  fp64            plain path: FP64 advection / reinit, FP64 Jacobi PCG;
  mixed           FP32 advection / reinit, FP32 Jacobi PCG refined in FP64;
  multigrid       FP32 advection / reinit, flexible CG with an algebraic
                  multigrid V-cycle (FP32 cycle, FP64 Krylov vectors);
  learned-jacobi  FP32 advection / reinit, flexible CG preconditioned by the
                  U-Net, the mixed Jacobi PCG taking over when it stalls;
  learned         the same with the V-cycle taking over (the headline).
The network is trained first, untimed, online on the device: 128^3 dam
breaks with seeded box extents and drop positions U[0.2, 0.8], a few Adam
steps after every step on pairs made from that step's own pressure system
(its (divergence, pressure) pair and (A e, e) pairs of smooth random e).
The headline value is --mode (default learned); the others are reported
beside it with their speed-up over fp64 and their volume drift against it.
After the timed steps each configuration runs on to --trajectory steps (the
workload's 300), timed per step, and the whole trajectory is summarised.

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
MFMA_BF16_PEAK_TFLOPS = 2500.0   # dense BF16 MFMA
# HBM bytes per launch of the dominant kernel at 256^3 from two rocprofv3 --pmc passes
# (FETCH_SIZE, WRITE_SIZE, scaled by a copy kernel's reading):
# profiles/dam_break_256_reinit_stage_traffic.md.  Used when --traffic is not given.
MEASURED_TRAFFIC = {("reinit_stage", 256): 2.824e8}
# Vector-ALU issue of the same kernel (SQ_INSTS_VALU against two wave64
# instructions per CU-cycle): profiles/dam_break_256_valu_issue.md.
MEASURED_VALU_ISSUE = {("reinit_stage", 256): 0.57}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# name: (pressure solve, advection precision, what a stalled learned solve continues with)
MODES = {"fp64": (capi.PCG_FP64, capi.FP64, 0), "mixed": (capi.PCG_MIXED, capi.FP32, 0),
         "learned": (capi.PCG_LEARNED, capi.FP32, capi.PCG_MULTIGRID), "learned-jacobi": (capi.PCG_LEARNED, capi.FP32, 0),
         "multigrid": (capi.PCG_MULTIGRID, capi.FP32, 0)}


# what each configuration's pressure solve is, as the JSON line states it
PRECONDITIONER = {
    "fp64": "FP64 Jacobi PCG",
    "mixed": "FP32 Jacobi PCG refined in FP64 (residual replacement, FP64 reductions)",
    "multigrid": "flexible CG (FP64) with an FP32 algebraic-multigrid V-cycle",
    "learned-jacobi": "flexible CG (FP64) with the U-Net cycle (Jacobi pre/post-smoothing, correction scaled to "
                      "minimise the A-norm error); once the A-norm error stops halving per iteration the mixed "
                      "Jacobi PCG finishes the solve",
    "learned": "flexible CG (FP64) with the U-Net cycle until the A-norm error stops halving per iteration "
               "(at least two applications), then the algebraic-multigrid V-cycle: the step's speed over FP64 "
               "comes from the V-cycle and the FP32 kernels, the network does not shorten the solve",
}


def build_scene(sim, g, args, rng=None):
    """The benchmark scene, or with `rng` a randomised one for training:
    dam-break box extents U[0.2, 0.8]; droplet positions seeded."""
    if args.scene == "droplet":
        return scenes.droplet(sim, g, seed=args.scene_seed if rng is None else int(rng.integers(1 << 30)))
    ext = (0.4, 0.6, 0.5) if rng is None else tuple(float(x) for x in rng.uniform(0.2, 0.8, size=3))
    scenes.dam_break(sim, g, ext)
    return ext


def train_network(args):
    """Network trained online on the simulator's own pressure systems
    (untimed): train_runs dam breaks at train_n^3 with box extents and a drop
    position U[0.2, 0.8] (seeded), train_pairs steps each; after every step
    train_adam Adam steps on pairs of that step's system."""
    import numpy as np
    rng = np.random.default_rng(args.train_seed)
    net = capi.Unet(seed=args.train_seed)
    g = scenes.grid(args.train_n, 3)
    opts = capi.step_opts(reinit_interval=args.reinit_interval, tol=args.tol, max_iterations=args.max_iterations,
                          precision=capi.FP32, mode=capi.PCG_MULTIGRID)
    t0 = time.time()
    losses, pairs = [], 0
    for run in range(args.train_runs):
        sim = capi.Liquid(3, g["res"], g["dx"], g["origin"], g["bw"])
        ext = tuple(float(x) for x in rng.uniform(0.2, 0.8, size=3))
        drop = tuple(float(x) for x in rng.uniform(0.2, 0.8, size=3))
        scenes.dam_break_with_drop(sim, g, ext, drop)
        sim.initialize(capi.step_opts(reinit_interval=args.reinit_interval, tol=args.tol, max_iterations=args.max_iterations))
        clk = FrameClock(g["dx"], args.cfl)
        for s in range(args.train_pairs):
            sim.step(clk.next_dt(sim.max_speed()), s, opts)
            h = sim.train_network(net, args.train_adam, lr=args.train_lr, operator_fraction=args.train_operator_fraction,
                                  seed=int(rng.integers(1 << 62)))
            losses.append(float(h.mean()))
            pairs += 1
        sim.close()
        if run % 5 == 4 or run + 1 == args.train_runs:
            log(f"[bench] training run {run + 1}/{args.train_runs}: box {ext}, drop {drop}; loss {np.mean(losses[-args.train_pairs:]):.3f}")
    k = max(1, len(losses) // 10)
    info = {"pairs": pairs, "grid": f"{args.train_n}^3", "runs": args.train_runs, "adam_steps": pairs * args.train_adam,
            "adam_steps_per_pair": args.train_adam, "operator_fraction": args.train_operator_fraction, "lr": args.train_lr,
            "loss_first": round(float(np.mean(losses[:k])), 4), "loss_last": round(float(np.mean(losses[-k:])), 4),
            "loss": "masked squared error of unit-RMS targets (1 = no better than 0)", "train_s": round(time.time() - t0, 1)}
    log(f"[bench] network trained: {info}")
    return net, info


def run_gpu(args, mode_name, net=None):
    g = scenes.grid(args.n, 3)
    mode, precision, fallback = MODES[mode_name]
    opts = capi.step_opts(reinit_interval=args.reinit_interval, tol=args.tol, max_iterations=args.max_iterations,
                          check_interval=args.check_interval, mode=mode, precision=precision, fallback_mode=fallback)
    t0 = time.time()
    sim = capi.Liquid(3, g["res"], g["dx"], g["origin"], g["bw"])
    if net is not None:
        sim.attach_preconditioner(net)
    build_scene(sim, g, args)
    init = capi.step_opts(reinit_interval=args.reinit_interval, tol=args.tol, max_iterations=args.max_iterations,
                          check_interval=args.check_interval, mode=capi.PCG_FP64)
    sim.initialize(init)
    v0 = sim.volume()
    log(f"[bench] {mode_name}: {args.n}^3 initialised in {time.time() - t0:.1f}s")
    clk = FrameClock(g["dx"], args.cfl)
    step = 0
    # each step returns the largest speed of the velocity it leaves (the next
    # step's CFL input), formed on the device at its end
    speed = sim.max_speed()
    for _ in range(args.warmup):
        speed = sim.step(clk.next_dt(speed), step, opts)["max_speed"]
        step += 1
    log(f"[bench] {args.warmup} warm-up steps done")
    sim.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        dt = clk.next_dt(speed)
        stats.append(sim.step(dt, step, opts))
        speed = stats[-1]["max_speed"]
        step += 1
        if k % 5 == 4:
            log(f"[bench] timed step {k + 1}/{args.steps}")
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    v1 = sim.volume()
    traj = None
    if args.trajectory > step:
        # on to the end of the workload's trajectory, each step timed on its own
        # (a step returns once its work is done)
        all_stats, wall = list(stats), []
        while step < args.trajectory:
            dt = clk.next_dt(speed)
            t1 = time.perf_counter()
            all_stats.append(sim.step(dt, step, opts))
            wall.append(time.perf_counter() - t1)
            speed = all_stats[-1]["max_speed"]
            step += 1
            if step % 50 == 0:
                log(f"[bench] {mode_name}: trajectory step {step}/{args.trajectory}")
        traj = (all_stats[len(stats):], wall, (sim.volume() - v0) / v0)
    sim.set_precision(precision)   # the kernels timed for the roofline are the ones this mode ran
    return g, sim, stats, elapsed, (v1 - v0) / v0, traj


def phase_roofline(stats, cells):
    """Algorithmic FP64 bytes per cell of each phase (one read and one write of
    every array it must touch, derived quantities on the fly) and the rate
    they imply at the measured phase times: advection a1-a7 433 B
    (216 + 48 + 169), WENO5 64 B per iteration (30 per reinitialisation),
    glue a13 266 B, projection set-up and gradient 160 B plus the solver's
    own count (hbm_bytes of the step's statistics: 136 B (FP64) or 68 B
    (FP32 inner) per unknown per Jacobi-PCG iteration; for the flexible
    solves their FCG vectors on the box or the compact rows, every V-cycle
    level's sweeps, restriction and prolongation, and the vector work around
    each network application, the network itself excluded (MFMA-bound,
    reported under `network`))."""
    K = len(stats)
    reinit = 30 * 64.0 * sum(s["reinitialized"] + s["restarted"] * (1 - s["reinitialized"]) for s in stats) / K
    solver = sum(s["pcg"]["hbm_bytes"] for s in stats) / K / cells
    phases = {"advect": (433.0, "t_advect_ms"), "reinit_check": (reinit, "t_restart_ms"), "extrap_enforce": (266.0, "t_extrap_ms"),
              "project": (160.0 + solver, "t_project_ms")}
    out = {}
    for name, (b, key) in phases.items():
        ms = sum(s[key] for s in stats) / K
        out[name] = {"ms": round(ms, 3), "bytes_per_cell": round(b, 1) if b else None,
                     "gbs": round(b * cells / ms / 1e6, 1) if b and ms > 0 else None,
                     "frac": round(b * cells / ms / 1e6 / HBM_PEAK_GBS, 4) if b and ms > 0 else None}
    return out


def pcg_summary(stats):
    K = len(stats)
    it = [s["pcg"]["iterations"] for s in stats]
    return {"mean": sum(it) / K, "min": min(it), "max": max(it),
            "outer_mean": sum(s["pcg"]["outer_iterations"] for s in stats) / K,
            "network_applications_mean": sum(s["pcg"]["preconditioner_applications"] for s in stats) / K,
            "fell_back": sum(s["pcg"]["fell_back"] for s in stats)}


def summary(stats, elapsed, cells, drift, unknown_fraction, traj):
    K = len(stats)
    out = {"value": round(cells * K / elapsed, 1), "ms_per_step": round(1000 * elapsed / K, 3),
           "unknown_fraction": round(unknown_fraction, 4),
           "phase_roofline": phase_roofline(stats, cells),
           "phases_ms": {k: round(sum(s[k] for s in stats) / K, 3) for k in ("t_advect_ms", "t_restart_ms", "t_project_ms", "t_extrap_ms", "t_total_ms")},
           "pcg_iterations": pcg_summary(stats),
           "pcg_residual_max": max(s["pcg"]["relative_residual"] for s in stats),
           "restarts": sum(s["restarted"] for s in stats),
           "volume_drift": drift}
    if traj:
        tstats, wall, tdrift = traj
        all_stats = stats + tstats
        n_all = len(all_stats)
        out["trajectory"] = {
            "steps_run": n_all, "steps_after_timed": len(tstats),
            "ms_per_step_after_timed": round(1000 * sum(wall) / len(wall), 3),
            "ms_per_step_all": round((1000 * elapsed + 1000 * sum(wall)) / n_all, 3),
            "phases_ms": {k: round(sum(s[k] for s in all_stats) / n_all, 3) for k in ("t_advect_ms", "t_restart_ms", "t_project_ms", "t_extrap_ms")},
            "project_roofline": phase_roofline(all_stats, cells)["project"],
            "pcg_iterations": pcg_summary(all_stats),
            "pcg_residual_max": max(s["pcg"]["relative_residual"] for s in all_stats),
            "restarts": sum(s["restarted"] for s in all_stats),
            "volume_drift": tdrift}
    return out


def roofline(sim, stats, args, mode_name):
    K = len(stats)
    # SpMV launches per step of the Jacobi-preconditioned solvers; the flexible
    # solves (learned, multigrid) run kernels of their own, not timed here
    flexible = MODES[mode_name][0] in (capi.PCG_LEARNED, capi.PCG_MULTIGRID)
    iters = 0 if flexible else sum(s["pcg"]["iterations"] + s["pcg"]["outer_iterations"] for s in stats) / K
    mixed = mode_name != "fp64"
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
    # the dominant kernel: the single-launch kernel with the most time per step
    # (advect_refmap is two record packs and the trace, advect_velocity one
    # launch per axis: reported above, not candidates)
    single = [k for k in kernels if k not in ("advect_refmap", "advect_velocity")]
    dom = max(single, key=lambda k: kernels[k]["ms_per_step"])
    d = kernels[dom]
    # a plain FP64 copy of one cell field: the bandwidth a streaming kernel reaches here
    ms, by = sim.time_kernel(capi.K_COPY, reps=args.kernel_reps)
    copy = {"avg_ms": round(ms, 5), "bytes": by, "gbs": round(by / ms / 1e6, 1)}
    return dom, kernels, {"kernel": dom, "bound": "hbm", "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(d["gbs"] / HBM_PEAK_GBS, 4),
                          "traffic": args.traffic if args.traffic is not None else MEASURED_TRAFFIC.get((dom, args.n)),
                          "valu_issue_frac": MEASURED_VALU_ISSUE.get((dom, args.n)),
                          "avg_launch_ms": round(d["avg_ms"], 5), "bytes_per_launch": d["bytes"],
                          "copy_kernel": copy}


def network_roofline(sim, net, res, reps=10):
    """One U-Net application on the bounding box the learned solve uses (the
    unknowns' box, two cells of context, sizes rounded up to multiples of 4),
    timed with HIP events over synthetic input, against the dense BF16 MFMA peak."""
    import numpy as np
    g = sim.get(capi.UNKNOWNS).reshape(res[2], res[1], res[0]) != 0
    dims = []
    for axis in (0, 1, 2):   # z, y, x
        idx = np.nonzero(g.any(axis=tuple(a for a in (0, 1, 2) if a != axis)))[0]
        dims.append(((int(idx[-1]) - int(idx[0]) + 5) + 3) // 4 * 4)
    ms, flops = net.time_forward(tuple(dims), reps=reps)
    tf = flops / ms / 1e9
    return {"box": [dims[2], dims[1], dims[0]], "ms_per_application": round(ms, 4), "gflop_per_application": round(flops / 1e9, 1),
            "bound": "mfma", "achieved": round(tf, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)}


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
    ap.add_argument("--scene", choices=["dam_break", "droplet"], default="dam_break")
    ap.add_argument("--scene-seed", type=int, default=19260817, help="droplet position seed")
    ap.add_argument("--mode", choices=list(MODES), default="learned", help="headline configuration")
    ap.add_argument("--also", default="fp64,mixed,multigrid,learned-jacobi", help="configurations run beside the headline one ('' for none)")
    ap.add_argument("--trajectory", type=int, default=300, help="run every configuration on to this many steps after the timed ones (0: stop)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iterations", type=int, default=20000)
    ap.add_argument("--check-interval", type=int, default=16)
    ap.add_argument("--reinit-interval", type=int, default=5)
    ap.add_argument("--cfl", type=float, default=1.0)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--train-n", type=int, default=128)
    ap.add_argument("--train-runs", type=int, default=20)
    ap.add_argument("--train-pairs", type=int, default=200, help="simulation steps (pressure systems) per training run")
    ap.add_argument("--train-adam", type=int, default=2, help="Adam steps after every simulation step")
    ap.add_argument("--train-operator-fraction", type=float, default=0.75)
    ap.add_argument("--train-lr", type=float, default=1e-3)
    ap.add_argument("--train-seed", type=int, default=19260817)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch of the dominant kernel measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                         "(scaled by the copy kernel's reading), reported as roofline.traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU decomposition is designed "
              "but not built", file=sys.stderr)
        sys.exit(2)

    cells = args.n ** 3
    order = [m for m in args.also.split(",") if m and m != args.mode] + [args.mode]
    results, net, train_info, network = {}, None, None, None
    for m in order:
        learned = MODES[m][0] == capi.PCG_LEARNED
        if learned and net is None:
            net, train_info = train_network(args)
        g, sim, stats, elapsed, drift, traj = run_gpu(args, m, net if learned else None)
        unknown_fraction = float((sim.get(capi.UNKNOWNS) != 0).sum()) / cells   # at the last step
        results[m] = summary(stats, elapsed, cells, drift, unknown_fraction, traj)
        if learned and network is None:
            network = network_roofline(sim, net, g["res"])
        if m == args.mode:
            dom, kernels, roof = roofline(sim, stats, args, m)
        sim.close()
    head = results[args.mode]
    plain = results.get("fp64")
    for m, r in results.items():
        if plain and m != "fp64":
            r["speedup_vs_fp64"] = round(r["value"] / plain["value"], 3)
            r["volume_drift_difference_vs_fp64"] = abs(r["volume_drift"] - plain["volume_drift"])
            if "trajectory" in r and "trajectory" in plain:
                t, tp = r["trajectory"], plain["trajectory"]
                t["speedup_vs_fp64"] = round(tp["ms_per_step_all"] / t["ms_per_step_all"], 3)
                t["volume_drift_difference_vs_fp64"] = abs(t["volume_drift"] - tp["volume_drift"])
    out = {
        "metric": METRIC, "value": head["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.mode == "fp64" else "f32+f64+bf16" if MODES[args.mode][0] == capi.PCG_LEARNED else "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"3D dam break full step, {args.n}^3 MAC grid (liquid box x<0.4, y<0.6, z<0.5, gravity -9.8, CFL {args.cfl:g})"
                                if args.scene == "dam_break" else
                                f"3D droplet into pool full step, {args.n}^3 MAC grid (pool y<0.3, sphere r=0.08 at seeded (x,0.7,z), "
                                f"u=(0,-2,0) in the drop, gravity -9.8, CFL {args.cfl:g}, seed {args.scene_seed})"),
                   "mode": args.mode, "preconditioner": PRECONDITIONER[args.mode], "tolerance": args.tol,
                   "reinit_interval": args.reinit_interval, "cells": cells},
        "phases_ms": head["phases_ms"],
        "pcg_iterations": head["pcg_iterations"],
        "pcg_residual_max": head["pcg_residual_max"],
        "speedup_vs_fp64": head.get("speedup_vs_fp64"),
        "volume_drift_difference_vs_fp64": head.get("volume_drift_difference_vs_fp64"),
        "trajectory": head.get("trajectory"),
        "modes": results,
        "network_training": train_info,
        "network": network,
        "kernels": {k: {kk: round(vv, 5) if isinstance(vv, float) else vv for kk, vv in v.items()} for k, v in kernels.items()},
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
