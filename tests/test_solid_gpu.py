"""GPU checks of the terms a box domain never reaches, and of the
fourth-order gradient, against hand values and the oracle.

* computeFirst<4> (restart and initialisation): exact for quartics inside,
  the clamped edge stencil by hand;
* the pressure row of a cell with partial face fractions and a moving solid
  (the (1 - w) u_solid term of Projector.cpp:58-59), by hand;
* Boundary::enforce on faces with 0.1 < fraction < 1, by hand;
* a dam break over a slanted solid floor with a moving-solid velocity, per
  step from identical inputs against the oracle at 1e-10 (FP64).
"""
import numpy as np
import pytest

from garm_amd import capi, scenes
from garm_amd.driver import FrameClock
from helpers import copy_state, make_gpu, make_oracle, rel_linf, say
from test_oracle_pins import (FIRST4_EDGE_X, FIRST4_EDGE_Y, hand_enforce, hand_projection, interior_mask,
                              quartic_field)

pytestmark = pytest.mark.gpu

BOUNDARY_FIELDS = [capi.NODE_DIST, capi.CELL_DIST, capi.FRACTION, capi.NORMAL, capi.SOLID_VELOCITY]


@pytest.mark.parametrize("dim", [2, 3])
def test_first4_exact_for_quartics_on_gpu(dim):
    say(f"computeFirst<4> quartic {dim}d")
    g = scenes.grid(14, dim)
    d = make_gpu(g)
    f, grad = quartic_field(scenes.cell_positions(g))
    d.set(capi.PHI, f)
    d.restart()
    got = d.get(capi.PHI_REF_GRAD).reshape(-1, dim)
    m = interior_mask(g, 2)
    assert np.max(np.abs(got[m] - grad[m])) <= 1e-10 * np.max(np.abs(grad[m]))


def test_first4_clamped_edges_on_gpu():
    say("computeFirst<4> clamped edges by hand")
    d = capi.Liquid(2, [6, 5], 1.0, [0.0, 0.0], 0)
    i, j = np.meshgrid(np.arange(6), np.arange(5))
    d.set(capi.PHI, (i ** 2 + 10.0 * j).ravel())
    d.restart()
    got = d.get(capi.PHI_REF_GRAD).reshape(5, 6, 2)
    for jj in range(5):
        assert np.allclose(got[jj, :, 0], FIRST4_EDGE_X, rtol=0, atol=1e-13)
    for ii in range(6):
        assert np.allclose(got[:, ii, 1], FIRST4_EDGE_Y, rtol=0, atol=1e-13)


def test_projection_partial_fractions_and_solid_velocity_by_hand_on_gpu():
    say("projection by hand: partial fractions, moving solid")
    d = capi.Liquid(2, [5, 5], 1.0, [0.0, 0.0], 1)
    p = hand_projection(d, "gpu")
    assert p == pytest.approx(3.35 / 4.3, rel=1e-14)


def test_enforce_partial_fraction_branch_by_hand_on_gpu():
    say("Boundary::enforce by hand: partial fractions")
    hand_enforce(make_gpu(scenes.grid(16, 2)))


def _solid_scene(g, o, d):
    """Dam break over the solid half-space below a slanted plane, with a
    solid velocity on every face (the oracle builds the boundary, the GPU
    takes it as uploads, as a host builder would)."""
    dim = g["dim"]
    o.solid_plane([0.0, 0.2, 0.0][:dim], [-0.25, 1.0, 0.15][:dim])
    vs = np.concatenate([np.full(len(scenes.face_positions(g, a)), (0.3, -0.2, 0.1)[a]) for a in range(dim)])
    o.set(capi.SOLID_VELOCITY, vs)
    for f in BOUNDARY_FIELDS:
        d.set(f, o.get(f))
    fr = o.get(capi.FRACTION)
    assert np.count_nonzero((fr > 0.1) & (fr < 1)) > 10 and np.count_nonzero((fr > 0) & (fr <= 0.1)) > 0
    for s in (o, d):
        scenes.dam_break(s, g)


@pytest.mark.parametrize("dim,n", [(2, 48), (3, 24)], ids=["2d-48", "3d-24"])
def test_solid_plane_scene_per_step(dim, n):
    g = scenes.grid(n, dim)
    o, d = make_oracle(g), make_gpu(g)
    _solid_scene(g, o, d)
    assert o.initialize(tol=1e-13, max_it=50000) == 0
    opts = capi.step_opts(reinit_interval=1, tol=1e-13, max_iterations=50000)
    d.initialize(opts)
    for f in (capi.PHI, capi.VELOCITY, capi.PRESSURE):
        e = rel_linf(d.get(f), o.get(f))
        assert e <= 1e-10, f"initialize: {capi.FIELD_NAMES[f]} {e:.2e}"
    clk = FrameClock(g["dx"], 1.0)
    for s in range(5):
        copy_state(o, d)
        dt = clk.next_dt(o.max_speed())
        say(f"solid scene step {s} {dim}d n={n} dt={dt:.3e}")
        if s == 3:   # a forced restart on both sides
            o.restart()
            d.restart()
        ro = o.step(dt, do_reinit=(s % 2 == 0), tol=1e-13, max_it=50000)
        rd = d.step(dt, 0 if s % 2 == 0 else 1, capi.step_opts(reinit_interval=2, tol=1e-13, max_iterations=50000))
        assert ro["rc"] == 0 and rd["pcg"]["unknowns"] == ro["unknowns"] and rd["restarted"] == ro["restarted"]
        for f in (capi.PHI, capi.PHI_GRAD, capi.VELOCITY, capi.PRESSURE, capi.REFMAP, capi.REFMAP_GRAD, capi.PHI_REF,
                  capi.ADV_REGION):
            a, b = d.get(f), o.get(f)
            if f == capi.ADV_REGION:
                assert np.array_equal(a, b)
            else:
                e = rel_linf(a, b)
                assert e <= 1e-10, f"step {s}: {capi.FIELD_NAMES[f]} {e:.2e}"
