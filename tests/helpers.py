"""Shared helpers for the parity tests: scene set-up on both sides, state
transfer and relative L-infinity comparison."""
import sys

import numpy as np

import oracle as O
from garm_amd import capi, scenes

# fields that make up the evolving state of LevelSetLiquid (LevelSetLiquid.h:27-66)
STATE_FIELDS = [capi.PHI, capi.PHI_GRAD, capi.REFMAP, capi.REFMAP_GRAD, capi.PHI_TMP, capi.PHI_ORI,
                capi.PHI_REF, capi.PHI_REF_GRAD, capi.ADV_REGION, capi.VELOCITY]


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    d = np.max(np.abs(a - b)) if a.size else 0.0
    return d / scale if scale > 0 else d


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def make_oracle(g):
    return O.Oracle(g["dim"], g["res"], g["dx"], g["origin"], g["bw"])


def make_gpu(g):
    return capi.Liquid(g["dim"], g["res"], g["dx"], g["origin"], g["bw"])


def copy_state(src, dst, fields=STATE_FIELDS):
    for f in fields:
        dst.set(f, src.get(f))


def evolved_oracle(g, steps=4, cfl=1.0, tol=1e-12, extents=(0.4, 0.6, 0.5), reinit_every=2):
    """A dam break run a few steps on the oracle so that J != I and u != 0."""
    from garm_amd.driver import FrameClock
    o = make_oracle(g)
    scenes.dam_break(o, g, extents)
    assert o.initialize(tol=tol) == 0
    clk = FrameClock(g["dx"], cfl)
    for s in range(steps):
        dt = clk.next_dt(o.max_speed())
        r = o.step(dt, do_reinit=(s % reinit_every == 0), tol=tol)
        assert r["rc"] == 0
    return o
