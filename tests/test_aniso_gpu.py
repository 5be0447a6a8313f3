"""A 3D dam break on an anisotropic 144 x 32 x 32 grid against the oracle.

Rows of 144 cells reach the kernel branches that isotropic parity grids of
24-40 cells never take (the WENO5 stage's unclamped x-row loads need at
least 131 cells in x; several x-waves per row in every kernel).  FP64: per
step from identical inputs, 1e-10 relative L-infinity for every evolved
field over 10 steps with reinitialisations and a forced restart.  FP32: the
same steps once more, the error of each field against the oracle stated and
bounded (phi within 5 dx of the interface).
"""
import numpy as np
import pytest

from garm_amd import capi, scenes
from garm_amd.driver import FrameClock
from helpers import copy_state, make_gpu, make_oracle, rel_linf, say

pytestmark = pytest.mark.gpu

RES = [144, 32, 32]
FIELDS = [capi.PHI, capi.PHI_GRAD, capi.VELOCITY, capi.PRESSURE, capi.REFMAP, capi.REFMAP_GRAD, capi.PHI_TMP,
          capi.PHI_ORI, capi.PHI_REF]


def _grid():
    bw = 2
    dx = 1.0 / (RES[1] - 2 * bw)     # y and z span [0, 1], x spans [0, 5]
    return dict(dim=3, res=list(RES), dx=dx, origin=[-bw * dx] * 3, bw=bw)


def _band(g, phi):
    return np.abs(phi) <= 5 * g["dx"]


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32], ids=["fp64", "fp32"])
def test_anisotropic_dam_break_per_step(precision):
    g = _grid()
    o, d = make_oracle(g), make_gpu(g)
    for s in (o, d):
        scenes.dam_break(s, g, extents=(1.2, 0.6, 0.5))
    assert o.initialize(tol=1e-13, max_it=50000) == 0
    clk = FrameClock(g["dx"], 1.0)
    worst = {capi.FIELD_NAMES[f]: 0.0 for f in FIELDS}
    for s in range(10):
        copy_state(o, d)
        dt = clk.next_dt(o.max_speed())
        reinit = s % 5 == 0
        if s == 7:
            o.restart()
            d.restart()
        say(f"aniso 144x32x32 {'fp64' if precision == capi.FP64 else 'fp32'} step {s} dt={dt:.3e} reinit={reinit}")
        mode = capi.PCG_FP64 if precision == capi.FP64 else capi.PCG_MIXED
        opts = capi.step_opts(reinit_interval=5, precision=precision, mode=mode, tol=1e-13, max_iterations=50000)
        ro = o.step(dt, do_reinit=reinit, tol=1e-13, max_it=50000)
        rd = d.step(dt, 0 if reinit else 1, opts)
        assert ro["rc"] == 0 and rd["pcg"]["unknowns"] == ro["unknowns"] and rd["restarted"] == ro["restarted"]
        band = _band(g, o.get(capi.PHI))
        for f in FIELDS:
            a, b = d.get(f), o.get(f)
            if precision == capi.FP32 and f in (capi.PHI, capi.PHI_TMP, capi.PHI_ORI):
                a, b = a[band], b[band]   # within 5 dx of the interface
            worst[capi.FIELD_NAMES[f]] = max(worst[capi.FIELD_NAMES[f]], rel_linf(a, b))
        assert np.array_equal(d.get(capi.ADV_REGION), o.get(capi.ADV_REGION))
    say("  worst rel Linf per field: " + ", ".join(f"{k} {v:.2e}" for k, v in worst.items()))
    if precision == capi.FP64:
        for k, v in worst.items():
            assert v <= 1e-10, f"{k}: {v:.2e}"
    else:
        # one FP32 step from FP64 inputs (phi within 5 dx of the interface).
        # Measured: phi 5.8e-3 (1.0e-3 absolute, 0.03 dx, after the 30 WENO5
        # iterations of step 5 at a point where the liquid's distance field
        # has a kink: the FP64 reinitialisation itself turns 1e-7 relative
        # noise on its input into 4.5e-5 there), grad phi 3.3e-4, u 2.9e-5,
        # p 1.7e-6, xi 5.9e-8, J 2.4e-4, phi before reinitialisation 1.2e-6.
        bounds = {"phi": 1e-2, "phi_grad": 1e-3, "velocity": 1e-4, "pressure": 1e-5, "refmap": 1e-6,
                  "refmap_grad": 1e-3, "phi_tmp": 1e-5, "phi_ori": 1e-5, "phi_ref": 1e-5}
        for k, v in worst.items():
            assert v <= bounds[k], f"{k}: {v:.2e} > {bounds[k]:.0e}"
