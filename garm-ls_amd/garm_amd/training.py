"""Training data of the learned pressure preconditioner.

Two kinds of pairs, both laid out on the network grid (the cell grid padded
with non-unknown cells to multiples of 4):
  - (divergence, pressure) of the simulator's own pressure solves: after each
    step of a device Liquid the right-hand side, the FP64 solution and the
    unknown indicator of its projection are read back (fields RHS, PRESSURE,
    UNKNOWNS);
  - (A x, x) for random smooth fields x on the unknowns of such a solve, with
    A the simulator's own pressure operator (garm_apply_pressure_system).
    Gravity-driven pressures are largely predictable from the liquid region
    alone; these pairs make the network follow its input.
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


def smooth_field(shape, rng, min_wavelength=3.0, max_wavelength=None):
    """Random field (x-fastest cell order of `shape` = (nx, ny, nz)) whose
    spectrum is band-limited to wavelengths >= min_wavelength cells with a
    random power-law slope, unit RMS."""
    nx, ny, nz = shape
    noise = rng.normal(size=(nz, ny, nx))
    kz = np.fft.fftfreq(nz)[:, None, None]
    ky = np.fft.fftfreq(ny)[None, :, None]
    kx = np.fft.rfftfreq(nx)[None, None, :]
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    k0 = 1.0 / (max_wavelength or max(shape))
    slope = rng.uniform(1.0, 3.0)
    filt = (1.0 + (k / k0) ** 2) ** (-slope / 2) * (k <= 1.0 / min_wavelength)
    f = np.fft.irfftn(np.fft.rfftn(noise) * filt, s=(nz, ny, nx))
    return f / max(np.sqrt(np.mean(f ** 2)), 1e-300)


def operator_pairs(sim, g, count, seed=0):
    """`count` pairs (A x, x) with random smooth x on the unknowns of the last
    projection of `sim`; returns stacked (rhs, solution, mask) in network layout."""
    rng = np.random.default_rng(seed)
    res = g["res"]
    shape = network_shape(res)
    unknowns = sim.get(capi.UNKNOWNS)
    mask = _on_network_grid(unknowns, res, shape, np.uint8)
    R, X, M = [], [], []
    for _ in range(count):
        x = smooth_field(res, rng).ravel() * unknowns
        b = sim.apply_pressure_system(x)
        R.append(_on_network_grid(b, res, shape, np.float32))
        X.append(_on_network_grid(x, res, shape, np.float32))
        M.append(mask)
    return np.stack(R), np.stack(X), np.stack(M)
