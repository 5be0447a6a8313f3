"""3D Enright deformation: advection and reinitialisation only.

Sphere of radius 0.15 at (0.35, 0.35, 0.35) in the unit domain, Enright
velocity with period T = 3 (garm_set_velocity_field), run to t = 3 at
CFL 0.5 (dt = 0.5 dx / 2, |u| <= 2), WENO5 reinitialisation every
`reinit_interval` steps (the reference: every step) and before every
restart.  Each step: velocity at the step's midpoint time
(the reference's test uses the end-of-step coefficient for the whole step,
TestAdvect3D.cpp:397; the midpoint keeps the RK2 trace second order),
GARM advection of xi and J, composition, reinitialisation, the distortion
check with restart, and the air extrapolation that rebuilds the advected
region (LevelSetLiquid.cpp:608-698 without projection, gravity and velocity
advection).

Run as a module for one JSON line comparing FP64 and FP32:
    python -m garm_amd.enright --n 256
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import capi, scenes


def run(n, precision=capi.FP64, t_end=3.0, period=3.0, cfl=0.5, reinit_interval=5, progress=None):
    g = scenes.grid(n, 3)
    sim = capi.Liquid(3, g["res"], g["dx"], g["origin"], g["bw"])
    sim.union_sphere([0.35, 0.35, 0.35], 0.15)
    sim.initialize(capi.step_opts(mode=capi.PCG_FP64, tol=1e-8))
    phi0 = sim.get(capi.PHI)
    v0, v0c = sim.volume(), sim.volume(corrected=True)
    sim.set_precision(precision)
    steps = int(math.ceil(t_end / (cfl * g["dx"] / 2.0)))
    dt = t_end / steps
    restarts = 0
    sim.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        sim.set_velocity_field(capi.VELOCITY_ENRIGHT, (s + 0.5) * dt, period)
        sim.advect_refmap(dt)
        sim.compose()
        reinit = s % reinit_interval == 0
        if reinit:
            sim.reinit(30, 0.9)
        if not sim.check_refmap():
            if not reinit:   # a restart captures a reinitialised level set, as garm_step does
                sim.reinit(30, 0.9)
            sim.restart()
            sim.compose()
            restarts += 1
        sim.extrapolate_air()
        if progress and s % 250 == 249:
            progress(f"[enright] n={n} {'fp32' if precision == capi.FP32 else 'fp64'} step {s + 1}/{steps} "
                     f"restarts {restarts} volume change {(sim.volume() - v0) / v0:+.4e}")
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    phi = sim.get(capi.PHI)
    out = {"steps": steps, "dt": dt, "restarts": restarts, "seconds": round(elapsed, 2),
           "volume_loss": (sim.volume() - v0) / v0, "volume_loss_corrected_order": (sim.volume(corrected=True) - v0c) / v0c}
    sim.close()
    return phi0, phi, out


def band_linf(a, b, ref, width):
    """max |a - b| over the cells where |ref| < width."""
    m = np.abs(ref) < width
    return float(np.max(np.abs(a - b)[m])) if m.any() else 0.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--reinit-interval", type=int, default=1)
    ap.add_argument("--fp64-only", action="store_true")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    dx = 1.0 / (args.n - 4)
    phi0, p64, r64 = run(args.n, capi.FP64, reinit_interval=args.reinit_interval, progress=log)
    if args.fp64_only:
        print(json.dumps({"fp64": r64}), flush=True)
        return
    _, p32, r32 = run(args.n, capi.FP32, reinit_interval=args.reinit_interval, progress=log)
    band = 3 * dx
    rep = {"workload": f"3D Enright deformation, {args.n}^3 MAC grid, sphere r=0.15 at (0.35,0.35,0.35), t=3, CFL 0.5",
           "fp64": r64, "fp32": r32,
           "phi_linf_fp32_vs_fp64_band3dx": band_linf(p32, p64, p64, band),
           "phi_linf_fp32_vs_fp64": float(np.max(np.abs(p32 - p64))),
           "phi_linf_fp64_vs_initial_band3dx": band_linf(p64, phi0, phi0, band),
           "phi_linf_fp32_vs_initial_band3dx": band_linf(p32, phi0, phi0, band),
           "dx": dx}
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
