"""The Jacobi PCG on a sparse bounding box of the unknowns, against the oracle.

Two blocks of liquid far apart along z (48 x 24 x 112 grid): the bounding
box of the unknowns is about twice the unknowns, and a third of its column
tiles hold none, so the 3D box solver launches only the tiles that hold
unknowns (the path a splash takes).  FP64: per step from identical inputs,
1e-10 relative L-infinity for the evolved fields over 6 steps with
reinitialisations; the mixed form: pressure within 1e-5 of the oracle's.
Both report covering fewer cells than the bounding box and more than the
unknowns (the tiles' cells, not the rows).
"""
import numpy as np
import pytest

from garm_amd import capi
from garm_amd.driver import FrameClock
from helpers import copy_state, make_gpu, make_oracle, rel_linf, say

pytestmark = pytest.mark.gpu

RES = [48, 24, 112]
FIELDS = [capi.PHI, capi.PHI_GRAD, capi.VELOCITY, capi.PRESSURE, capi.REFMAP, capi.REFMAP_GRAD]


def _grid():
    bw = 2
    dx = 1.0 / (RES[1] - 2 * bw)     # y spans [0, 1], x [0, 2.2], z [0, 5.4]
    return dict(dim=3, res=list(RES), dx=dx, origin=[-bw * dx] * 3, bw=bw)


def _blocks(sim):
    sim.union_box([0.0, 0.0, 0.0], [2.0, 0.5, 1.0])
    sim.union_box([0.0, 0.0, 3.5], [2.0, 0.5, 1.9])


def _box_cells(sim):
    idx = np.nonzero(sim.get(capi.UNKNOWNS).reshape(RES[::-1]) != 0)
    return int(np.prod([int(a.max()) - int(a.min()) + 1 for a in idx]))


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32], ids=["fp64", "mixed"])
def test_sparse_box_per_step(precision):
    g = _grid()
    o, d = make_oracle(g), make_gpu(g)
    for s in (o, d):
        _blocks(s)
    assert o.initialize(tol=1e-13, max_it=50000) == 0
    clk = FrameClock(g["dx"], 1.0)
    worst = {capi.FIELD_NAMES[f]: 0.0 for f in FIELDS}
    mode = capi.PCG_FP64 if precision == capi.FP64 else capi.PCG_MIXED
    for s in range(6):
        copy_state(o, d)
        dt = clk.next_dt(o.max_speed())
        reinit = s % 5 == 0
        opts = capi.step_opts(reinit_interval=5, precision=precision, mode=mode, tol=1e-13, max_iterations=50000)
        ro = o.step(dt, do_reinit=reinit, tol=1e-13, max_it=50000)
        rd = d.step(dt, 0 if reinit else 1, opts)
        pc = rd["pcg"]
        box = _box_cells(d)
        say(f"sparse box {'fp64' if precision == capi.FP64 else 'mixed'} step {s}: {pc['unknowns']} unknowns, "
            f"box {box} cells, solver covers {pc['solver_cells']}, {pc['iterations']} iterations")
        assert ro["rc"] == 0 and pc["unknowns"] == ro["unknowns"]
        assert pc["unknowns"] < pc["solver_cells"] < box and box > 1.5 * pc["unknowns"]
        for f in FIELDS:
            worst[capi.FIELD_NAMES[f]] = max(worst[capi.FIELD_NAMES[f]], rel_linf(d.get(f), o.get(f)))
        if precision == capi.FP32:
            break   # one mixed solve from the oracle's state: the pressure is the point
    say("  worst rel Linf per field: " + ", ".join(f"{k} {v:.2e}" for k, v in worst.items()))
    if precision == capi.FP64:
        for k, v in worst.items():
            assert v <= 1e-10, f"{k}: {v:.2e}"
    else:
        assert worst["pressure"] <= 1e-5, f"pressure: {worst['pressure']:.2e}"
