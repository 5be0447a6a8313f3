"""Host-side training data of the learned pressure preconditioner:
(divergence, pressure) pairs of the simulator's own pressure solves, read
back after each step of a device Liquid (fields RHS, PRESSURE, UNKNOWNS) and
laid out on the network grid (the cell grid padded with non-unknown cells to
multiples of 4), for garm_unet_train.  Training without host copies, on
pairs made on the device from each step's pressure system, is
Liquid.train_network (garm_unet_train_on_system).
"""
import numpy as np

from . import capi
from .driver import FrameClock


def network_shape(res):
    """(nz, ny, nx) of the network grid that holds a 3D cell grid `res`."""
    return tuple(((int(n) + 3) // 4) * 4 for n in (res[2], res[1], res[0]))


def _on_network_grid(a, res, shape, dtype):
    out = np.zeros(shape, dtype=dtype)
    nx, ny, nz = res
    out[:nz, :ny, :nx] = a.reshape(nz, ny, nx)
    return out


def read_pair(sim, res):
    """(rhs, pressure, mask) of the last projection of `sim`, network layout."""
    shape = network_shape(res)
    rhs = _on_network_grid(sim.get(capi.RHS), res, shape, np.float32)
    p = _on_network_grid(sim.get(capi.PRESSURE), res, shape, np.float32)
    m = _on_network_grid(sim.get(capi.UNKNOWNS), res, shape, np.uint8)
    return rhs, p, m


def collect_pairs(sim, g, steps, cfl=1.0, tol=1e-8, reinit_interval=1, every=1, first_step=0):
    """Advance a 3D device Liquid `steps` steps with FP64 PCG at `tol` and
    return the stacked (rhs, pressure, mask) of every `every`-th step."""
    if g["dim"] != 3:
        raise ValueError("the network is three-dimensional")
    opts = capi.step_opts(reinit_interval=reinit_interval, mode=capi.PCG_FP64, tol=tol, max_iterations=20000)
    clk = FrameClock(g["dx"], cfl)
    R, P, M = [], [], []
    for s in range(first_step, first_step + steps):
        dt = clk.next_dt(sim.max_speed())
        sim.step(dt, s, opts)
        if (s - first_step) % every == 0:
            r, p, m = read_pair(sim, g["res"])
            if m.any():
                R.append(r)
                P.append(p)
                M.append(m)
    return np.stack(R), np.stack(P), np.stack(M)
