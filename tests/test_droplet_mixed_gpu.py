"""The droplet-into-pool configuration with every lower-precision part on
(FP32 advection and reinitialisation, the learned preconditioner inside the
flexible CG with the mixed Jacobi solver and with the V-cycle as fallback)
against the plain FP64 GPU path on the same time steps, at 64^3 over 80
steps: every pressure solve of the combined form reaches the 1e-6 residual
measured in FP64, and its liquid volume drift stays within 0.1% of the
plain path's.  The network is trained online on the run's own first 16
pressure systems.
"""
import numpy as np
import pytest

from garm_amd import capi, scenes
from garm_amd.driver import FrameClock
from helpers import make_gpu, say

pytestmark = pytest.mark.gpu

N, STEPS, TRAIN = 64, 80, 16


def _scene(g):
    d = make_gpu(g)
    scenes.droplet(d, g, seed=19260817)
    d.initialize(capi.step_opts(tol=1e-10, max_iterations=20000))
    return d


@pytest.fixture(scope="module")
def plain():
    g = scenes.grid(N, 3)
    d = _scene(g)
    v0 = d.volume()
    clk = FrameClock(g["dx"], 1.0)
    opts = capi.step_opts(reinit_interval=5, tol=1e-6, max_iterations=20000)
    dts, vols = [], []
    for s in range(STEPS):
        dt = clk.next_dt(d.max_speed())
        r = d.step(dt, s, opts)
        assert r["pcg"]["relative_residual"] <= 1e-6
        dts.append(dt)
        vols.append(d.volume() / v0 - 1)
    say(f"droplet {N}^3 FP64: volume drift after {STEPS} steps {vols[-1]:+.3e}")
    return g, dts, np.array(vols)


@pytest.mark.parametrize("fallback", [capi.PCG_MIXED, capi.PCG_MULTIGRID], ids=["mixed-fallback", "multigrid-fallback"])
def test_fp32_learned_droplet_against_fp64(plain, fallback):
    g, dts, vols64 = plain
    d = _scene(g)
    v0 = d.volume()
    net = capi.Unet(seed=5)
    train = capi.step_opts(reinit_interval=5, precision=capi.FP32, mode=capi.PCG_MIXED, tol=1e-6, max_iterations=20000)
    opts = capi.step_opts(reinit_interval=5, precision=capi.FP32, mode=capi.PCG_LEARNED, tol=1e-6, max_iterations=20000,
                          fallback_mode=fallback)
    vols, apps, res = [], 0, []
    for s, dt in enumerate(dts):
        if s < TRAIN:
            r = d.step(dt, s, train)
            d.train_network(net, 20, lr=1e-3, seed=s)
            if s == TRAIN - 1:
                d.attach_preconditioner(net)
        else:
            r = d.step(dt, s, opts)
            apps += r["pcg"]["preconditioner_applications"]
        res.append(r["pcg"]["relative_residual"])
        vols.append(d.volume() / v0 - 1)
    vols = np.array(vols)
    diff = np.max(np.abs(vols - vols64))
    say(f"  FP32 + learned ({'mixed' if fallback == capi.PCG_MIXED else 'V-cycle'} fallback): worst residual {max(res):.2e}, "
        f"{apps} network applications, drift {vols[-1]:+.3e} vs FP64 {vols64[-1]:+.3e} (largest difference {diff:.2e})")
    assert max(res) <= 1e-6
    assert apps >= STEPS - TRAIN
    assert diff <= 1e-3
