"""Surface tension on the GPU (garm_apply_surface_tension, and garm_step with
surface_tension set) against the oracle's restatement, on buildCase1's
liquid column (LevelSetLiquidBuilder.cpp:67-87: box (0.4, 0.4, 1) in a
(2, 1.2, 1) domain, surface tension 30 over density 1e3, band 5.5 cells):

* from identical evolved states, one projection then the surface tension --
  velocity-only air extrapolation, boundary enforcement, the capillary solve
  on the interface faces, the reprojection -- the same interface faces and
  band, and the velocity and pressure within 1e-11 when the oracle solves
  both systems as the GPU does (Jacobi PCG from zero, to 1e-13); with the
  reference's solvers (IC(0) CG from the velocity guess, the reprojection
  from the projection's pressure) the same to what the tolerance leaves
  (the capillary diagonal 1/delta spans 0.15 to 3e5, so a 1e-13 residual
  still allows 1e-9 of the velocity's scale on the low-diagonal rows);
* whole steps with surface tension, per step from identical inputs;
* the reference's tolerances (1e-5 for both solves) with FP32 advection and
  the mixed pressure solve over 40 steps.
"""
import numpy as np
import pytest

import oracle as O
from garm_amd import capi, scenes
from garm_amd.driver import FrameClock
from helpers import copy_state, make_gpu, make_oracle, rel_linf, say

pytestmark = pytest.mark.gpu

CASES = [(2, 40), (3, 24)]
IDS = ["2d-40", "3d-24"]
TOL = 1e-13


def _oracle(dim, scale, steps, gpu_solvers=False):
    g = scenes.capillary_grid(scale, dim)
    o = make_oracle(g)
    scenes.capillary_column(o, g)
    assert o.initialize(tol=1e-12) == 0
    o.set_surface_tension(True, coeff=scenes.CAPILLARY_COEFF, delta_min=scenes.CAPILLARY_BAND, delta_max=scenes.CAPILLARY_BAND,
                          tol=TOL, max_it=50000)
    if gpu_solvers:
        o.set_surface_tension_solvers(O.PRECOND_JACOBI, capillary_guess=False, reproject_warm=False)
    clk = FrameClock(g["dx"], 1.0)
    for s in range(steps):
        r = o.step(clk.next_dt(o.max_speed()), do_reinit=(s % 2 == 0), tol=1e-12, max_it=50000)
        assert r["rc"] == 0
    return g, o, clk


def _faces_next_to_unknowns(g, unknown):
    """Faces whose adjacent cells include a pressure unknown (the faces
    Projector::applyPressureGradient can change, Projector.cpp:91-123)."""
    dim, res = g["dim"], g["res"]
    u = unknown.reshape(res[::-1])
    out = []
    for a in range(dim):
        ax = dim - 1 - a                      # numpy axis of grid axis a
        pad = [(0, 0)] * dim
        pad[ax] = (1, 1)
        p = np.pad(u, pad)
        lo = [slice(None)] * dim
        hi = [slice(None)] * dim
        lo[ax] = slice(0, res[a] + 1)
        hi[ax] = slice(1, res[a] + 2)
        out.append((p[tuple(lo)] | p[tuple(hi)]).ravel())
    return np.concatenate(out)


def _opts():
    return capi.surface_tension_opts(coeff=scenes.CAPILLARY_COEFF, delta_min=scenes.CAPILLARY_BAND,
                                     delta_max=scenes.CAPILLARY_BAND, tolerance=TOL, max_iterations=50000)


@pytest.mark.parametrize("dim,scale", CASES, ids=IDS)
def test_surface_tension_matches_oracle(dim, scale):
    g, o, clk = _oracle(dim, scale, 3)
    d = make_gpu(g)
    copy_state(o, d)
    d.elapsed_time = o.elapsed_time
    say(f"surface tension {dim}d scale {scale}: res {g['res']}")
    ref = make_oracle(g)
    copy_state(o, ref)
    ref.elapsed_time = o.elapsed_time
    ref.set_surface_tension(True, coeff=scenes.CAPILLARY_COEFF, delta_min=scenes.CAPILLARY_BAND, delta_max=scenes.CAPILLARY_BAND,
                            tol=TOL, max_it=50000)
    o.set_surface_tension_solvers(O.PRECOND_JACOBI, capillary_guess=False, reproject_warm=False)
    for s in (o, ref):
        assert s.project(tol=1e-12, max_it=50000)["rc"] == 0
    d.project(mode=capi.PCG_FP64, tol=1e-12, max_iterations=50000)
    assert rel_linf(d.get(capi.VELOCITY), o.get(capi.VELOCITY)) <= 1e-10
    dt = clk.next_dt(o.max_speed())
    rr = ref.apply_surface_tension(dt, tol=1e-13, max_it=50000)
    ro = o.apply_surface_tension(dt, tol=1e-13, max_it=50000)
    rd = d.apply_surface_tension(dt, _opts(), mode=capi.PCG_FP64, tol=1e-13, max_iterations=50000)
    say(f"  capillary faces {rd['capillary']['unknowns']} (oracle {ro['capillary']['faces']}), "
        f"iterations {rd['capillary']['iterations']} (oracle IC(0) {ro['capillary']['iterations']}), "
        f"band {rd['band']}")
    assert ro["rc"] == 0
    assert rd["capillary"]["unknowns"] == ro["capillary"]["faces"] > 0
    assert rd["band"] == ro["capillary"]["band"]
    assert rd["capillary"]["relative_residual"] <= TOL
    # the reprojection moves only faces next to a pressure unknown; elsewhere
    # the velocity is the capillary solve's (and the extrapolation's) alone
    uo, ud = o.get(capi.VELOCITY), d.get(capi.VELOCITY)
    touched = _faces_next_to_unknowns(g, d.get(capi.UNKNOWNS) != 0)
    scale_u = np.max(np.abs(uo))
    e_cap = np.max(np.abs(ud - uo)[~touched]) / scale_u
    e_all = rel_linf(ud, uo)
    e_p = rel_linf(d.get(capi.PRESSURE), o.get(capi.PRESSURE))
    say(f"  velocity: {e_all:.2e} ({e_cap:.2e} away from the pressure unknowns), pressure {e_p:.2e}")
    assert e_cap <= 1e-11 and e_all <= 1e-11 and e_p <= 1e-11
    # the reference's solvers reach the same fields to their tolerance
    assert rr["rc"] == 0 and rr["capillary"]["faces"] == ro["capillary"]["faces"]
    e_ref_u = rel_linf(ud, ref.get(capi.VELOCITY))
    e_ref_p = rel_linf(d.get(capi.PRESSURE), ref.get(capi.PRESSURE))
    say(f"  against the reference's solvers: velocity {e_ref_u:.2e}, pressure {e_ref_p:.2e}")
    assert e_ref_u <= 1e-7 and e_ref_p <= 1e-6


@pytest.mark.parametrize("dim,scale", CASES, ids=IDS)
def test_steps_with_surface_tension(dim, scale):
    g, o, clk = _oracle(dim, scale, 1, gpu_solvers=True)
    d = make_gpu(g)
    st = _opts()
    opts = capi.step_opts(reinit_interval=1, tol=1e-12, max_iterations=50000, surface_tension=st)
    for s in range(5):
        copy_state(o, d)
        d.elapsed_time = o.elapsed_time
        dt = clk.next_dt(o.max_speed())
        ro = o.step(dt, do_reinit=True, tol=1e-12, max_it=50000)
        rd = d.step(dt, s, opts)
        assert ro["rc"] == 0 and rd["restarted"] == ro["restarted"]
        assert rd["surface_tension"]["capillary"]["unknowns"] == o.capillary_stats()["faces"]
        assert d.elapsed_time == o.elapsed_time
        for f in (capi.PHI, capi.VELOCITY, capi.PRESSURE, capi.REFMAP, capi.ADV_REGION):
            e = rel_linf(d.get(f), o.get(f))
            assert e <= 1e-8, f"step {s}: {capi.FIELD_NAMES[f]} {e:.2e}"
    say(f"  {dim}d: 5 steps with surface tension agree")


def test_reference_tolerances_fp32_mixed():
    g = scenes.capillary_grid(64, 2)
    d = make_gpu(g)
    scenes.capillary_column(d, g)
    d.initialize(capi.step_opts(tol=1e-10, max_iterations=20000))
    st = capi.surface_tension_opts(coeff=scenes.CAPILLARY_COEFF, delta_min=scenes.CAPILLARY_BAND, delta_max=scenes.CAPILLARY_BAND)
    opts = capi.step_opts(reinit_interval=5, precision=capi.FP32, mode=capi.PCG_MIXED, tol=1e-5, surface_tension=st)
    clk = FrameClock(g["dx"], 1.0)
    v0 = d.volume()
    iters = []
    for s in range(40):
        r = d.step(clk.next_dt(d.max_speed()), s, opts)
        c = r["surface_tension"]["capillary"]
        assert c["unknowns"] > 0 and c["relative_residual"] <= 1e-5
        assert r["surface_tension"]["reproject"]["relative_residual"] <= 1e-5
        iters.append(c["iterations"])
    drift = abs(d.volume() / v0 - 1)
    say(f"  40 steps at 2d scale 64: capillary iterations {min(iters)}..{max(iters)}, volume drift {drift:.2e}")
    assert np.all(np.isfinite(d.get(capi.VELOCITY)))
    assert drift < 0.05
