"""Hybrid-Newton restart on the GPU (garm_restart_newton, the default
build's LevelSetLiquid::restart with enable-fmm=n) against the oracle's
restatement: from identical evolved states, the reference level set, its
gradient and the reset reference map within 1e-10; steps that restart this
way, per step from identical inputs; and the plane pin of the oracle
(Newton lands on the exact closest point) on the GPU itself.
"""
import numpy as np
import pytest

from garm_amd import capi, scenes
from garm_amd.driver import FrameClock
from helpers import copy_state, evolved_oracle, make_gpu, make_oracle, rel_linf, say
from test_oracle_pins import interior_mask

pytestmark = pytest.mark.gpu

CASES = [(2, 48), (3, 24)]
IDS = ["2d-48", "3d-24"]


@pytest.mark.parametrize("dim,n", CASES, ids=IDS)
def test_newton_restart_matches_oracle(dim, n):
    g = scenes.grid(n, dim)
    o = evolved_oracle(g, steps=6)
    d = make_gpu(g)
    copy_state(o, d)
    say(f"Newton restart {dim}d n={n}")
    o.restart_newton()
    d.restart_newton()
    for f in (capi.PHI_REF, capi.PHI_REF_GRAD, capi.REFMAP, capi.REFMAP_GRAD):
        e = rel_linf(d.get(f), o.get(f))
        assert e <= 1e-10, f"{capi.FIELD_NAMES[f]}: {e:.2e}"
    # the composition from the new reference level set agrees too
    o.compose()
    d.compose()
    for f in (capi.PHI, capi.PHI_GRAD):
        assert rel_linf(d.get(f), o.get(f)) <= 1e-10


@pytest.mark.parametrize("dim,n", CASES, ids=IDS)
def test_steps_with_newton_restarts(dim, n):
    """Steps in restart mode NEWTON from identical inputs until the distortion
    check has restarted at least once on both sides."""
    g = scenes.grid(n, dim)
    o, d = make_oracle(g), make_gpu(g)
    for s in (o, d):
        scenes.dam_break(s, g)
    assert o.initialize(tol=1e-12) == 0
    o.set_restart_newton(True)
    opts = capi.step_opts(reinit_interval=1, tol=1e-12, max_iterations=50000, restart_mode=capi.RESTART_NEWTON)
    clk = FrameClock(g["dx"], 1.0)
    restarts = 0
    for s in range(60):
        copy_state(o, d)
        dt = clk.next_dt(o.max_speed())
        ro = o.step(dt, do_reinit=True, tol=1e-12, max_it=50000)
        rd = d.step(dt, s, opts)
        assert ro["rc"] == 0 and rd["restarted"] == ro["restarted"]
        for f in (capi.PHI, capi.PHI_REF, capi.PHI_REF_GRAD, capi.VELOCITY, capi.PRESSURE):
            e = rel_linf(d.get(f), o.get(f))
            assert e <= 1e-10, f"step {s}: {capi.FIELD_NAMES[f]} {e:.2e}"
        restarts += ro["restarted"]
        if restarts >= 1 and s >= 8:
            break
    say(f"  {dim}d n={n}: {s + 1} steps, {restarts} Newton restarts")
    assert restarts >= 1


@pytest.mark.parametrize("dim", [2, 3])
def test_newton_plane_pin_on_gpu(dim):
    n = 40 if dim == 2 else 24
    g = scenes.grid(n, dim)
    d = make_gpu(g)
    x = scenes.cell_positions(g)
    nrm = np.array([0.6, 0.8] if dim == 2 else [0.48, 0.64, 0.6])
    sd = x @ nrm - 0.45
    d.set(capi.PHI, sd)
    d.set(capi.PHI_ORI, 2 * sd)
    d.set(capi.PHI_GRAD, np.tile(2 * nrm, (len(sd), 1)).ravel())
    d.restart_newton()
    pr, gr = d.get(capi.PHI_REF), d.get(capi.PHI_REF_GRAD).reshape(-1, dim)
    band = (np.abs(sd) <= 0.6 * g["dx"]) & interior_mask(g, 4)
    assert np.max(np.abs(pr[band] - sd[band])) <= 1e-13
    assert np.max(np.abs(gr[band] - nrm)) <= 1e-13
