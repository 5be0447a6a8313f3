"""The multigrid V-cycle at a size where level 0 holds more than 2^20 cells,
so its sweeps run on the large-tile kernel (32 x 8 x 8 cells per
workgroup) that the small parity cases never reach.  A 128^3 dam break
whose liquid spans the domain in x and z (level 0: 128 x 77 x 128 cells):
flexible CG with the V-cycle reaches a 1e-12 FP64 residual and the pressure
of the plain FP64 Jacobi PCG, in a tenth of its iterations or fewer.  The
reference answer is the GPU's FP64 PCG (the parity tests pin it to the
oracle; a CPU solve at this size takes minutes)."""
import numpy as np
import pytest

from garm_amd import capi, scenes
from helpers import make_gpu, rel_linf, say

pytestmark = pytest.mark.gpu


def _scene(g):
    d = make_gpu(g)
    scenes.dam_break(d, g, extents=(1.0, 0.6, 1.0))
    d.initialize(capi.step_opts(mode=capi.PCG_FP64, tol=1e-10, max_iterations=20000))
    d.gravity(0.02)
    return d


def test_multigrid_large_level0_matches_fp64_pcg():
    g = scenes.grid(128, 3)
    say("multigrid 3d n=128, liquid across x and z")
    ref, mg = _scene(g), _scene(g)
    rj = ref.project(mode=capi.PCG_FP64, tol=1e-12, max_iterations=50000)
    rm = mg.project(mode=capi.PCG_MULTIGRID, tol=1e-12, max_iterations=50000)
    e = rel_linf(mg.get(capi.PRESSURE), ref.get(capi.PRESSURE))
    say(f"  unknowns {rm['unknowns']}; iterations: multigrid {rm['iterations']}, Jacobi {rj['iterations']}; "
        f"residual {rm['relative_residual']:.2e}; pressure rel Linf {e:.2e}")
    assert rm["unknowns"] == rj["unknowns"] and rm["unknowns"] > 1 << 20
    assert rm["relative_residual"] <= 1e-12 and rm["fell_back"] == 0
    assert rm["iterations"] * 10 <= rj["iterations"]
    # both solves stop at a 1e-12 relative residual; the condition number of
    # the 128^3 system bounds their pressure difference far below this
    assert e <= 1e-7
    ev = rel_linf(mg.get(capi.VELOCITY), ref.get(capi.VELOCITY))
    assert ev <= 1e-7, ev


def test_multigrid_sparse_box_matches_fp64_pcg():
    """A droplet above a pool: the bounding box of the unknowns holds about
    twice as many cells as there are unknowns, so the flexible solve keeps its
    vectors compact (the dam breaks above run them on the box)."""
    g = scenes.grid(64, 3)
    say("multigrid 3d n=64, droplet above a pool")
    sims = []
    for _ in range(2):
        d = make_gpu(g)
        scenes.droplet(d, g)
        d.initialize(capi.step_opts(mode=capi.PCG_FP64, tol=1e-10, max_iterations=20000))
        d.gravity(0.02)
        sims.append(d)
    ref, mg = sims
    rj = ref.project(mode=capi.PCG_FP64, tol=1e-12, max_iterations=50000)
    rm = mg.project(mode=capi.PCG_MULTIGRID, tol=1e-12, max_iterations=50000)
    e = rel_linf(mg.get(capi.PRESSURE), ref.get(capi.PRESSURE))
    say(f"  unknowns {rm['unknowns']}; iterations: multigrid {rm['iterations']}, Jacobi {rj['iterations']}; "
        f"residual {rm['relative_residual']:.2e}; pressure rel Linf {e:.2e}")
    assert rm["unknowns"] == rj["unknowns"]
    assert rm["relative_residual"] <= 1e-12 and rm["fell_back"] == 0
    assert rm["iterations"] * 5 <= rj["iterations"]
    assert e <= 1e-7
    ev = rel_linf(mg.get(capi.VELOCITY), ref.get(capi.VELOCITY))
    assert ev <= 1e-7, ev


@pytest.mark.parametrize("mode", ["multigrid", "learned"])
def test_flexible_solvers_zero_right_hand_side(mode):
    """A liquid at rest without gravity: the right-hand side is exactly zero.
    The flexible solvers enqueue their first iteration before they read |b|^2
    back, so that iteration runs on b = 0 (its step lengths are 0 where p.Ap
    is) and the solve must still return p = 0, no iterations and a zero
    residual, and leave the velocity as it was.  The learned form runs an
    untrained network, whose output on a zero residual is not zero (its
    second input channel is the mask of the unknowns)."""
    g = scenes.grid(32, 3)
    say(f"zero right-hand side, {mode}, 3d n=32")
    d = make_gpu(g)
    scenes.dam_break(d, g, extents=(1.0, 0.5, 1.0))
    d.initialize(capi.step_opts(mode=capi.PCG_FP64, tol=1e-10, max_iterations=20000))
    u0 = d.get(capi.VELOCITY)
    assert not np.any(u0)
    kw = {}
    if mode == "learned":
        net = capi.Unet(seed=3)
        d.attach_preconditioner(net)
        kw = dict(fallback_mode=capi.PCG_MULTIGRID)
    r = d.project(mode=capi.PCG_MULTIGRID if mode == "multigrid" else capi.PCG_LEARNED, tol=1e-8, max_iterations=100, **kw)
    say(f"  unknowns {r['unknowns']}, iterations {r['iterations']}, residual {r['relative_residual']}")
    assert r["unknowns"] > 0 and r["iterations"] == 0 and r["relative_residual"] == 0 and r["fell_back"] == 0
    assert not np.any(d.get(capi.PRESSURE))
    assert np.array_equal(d.get(capi.VELOCITY), u0)
