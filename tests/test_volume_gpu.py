"""Volume measurement on the GPU (garm_volume) against the oracle.

Random phi_tmp fields reach all 256 corner patterns of the 3D fraction and
both diagonal rules of the 2D contour; the GPU sums in a different order,
so the bar is 1e-12 relative.  A free trajectory of the plain FP64 path
must keep the oracle's volume drift within 0.1% of the initial volume.
The ambiguous corner configurations are also checked on the GPU directly
against the closed forms of the marching-cubes geometry (mc_geometry.py),
one cube at a time, since the GPU's case table and the oracle's are two
transcriptions of the same data.
"""
import numpy as np
import pytest

import oracle as O
from garm_amd import capi, scenes
from garm_amd.driver import FrameClock
import mc_geometry as G
from helpers import copy_state, make_gpu, make_oracle, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,n", [(3, 24), (2, 48)], ids=["3d-24", "2d-48"])
@pytest.mark.parametrize("corrected", [False, True], ids=["reference-order", "corrected"])
def test_random_field_volume_matches_oracle(dim, n, corrected):
    say(f"volume random field {dim}d n={n} corrected={corrected}")
    g = scenes.grid(n, dim)
    o, d = make_oracle(g), make_gpu(g)
    rng = np.random.default_rng(n + dim)
    phi = rng.uniform(-1.0, 1.0, size=o.count(capi.PHI_TMP))
    o.set(capi.PHI_TMP, phi)
    d.set(capi.PHI_TMP, phi)
    vo, vd = o.volume(corrected), d.volume(corrected)
    say(f"  oracle {vo:.15e} gpu {vd:.15e}")
    assert abs(vd - vo) <= 1e-12 * abs(vo)
    if dim == 3:
        nx = g["res"][0]
        lex = phi.reshape(n, n, n)
        patterns = set()
        cube = np.stack([lex[k:k + n - 1, j:j + n - 1, i:i + n - 1] < 0 for k in (0, 1) for j in (0, 1) for i in (0, 1)])
        idx = np.zeros(cube.shape[1:], dtype=np.int64)
        for b in range(8):
            idx |= cube[b].astype(np.int64) << b
        patterns = np.unique(idx)
        assert len(patterns) == 256 and nx == n


@pytest.mark.parametrize("dim,n,steps", [(3, 24, 8), (2, 48, 12)], ids=["3d-24", "2d-48"])
def test_volume_drift_of_free_trajectory_matches_oracle(dim, n, steps):
    say(f"volume drift {dim}d n={n} steps={steps}")
    g = scenes.grid(n, dim)
    o, d = make_oracle(g), make_gpu(g)
    for s in (o, d):
        scenes.dam_break(s, g)
    assert o.initialize(tol=1e-12) == 0
    copy_state(o, d)
    d.set(capi.PHI_TMP, o.get(capi.PHI_TMP))
    v0o, v0d = o.volume(), d.volume()
    assert abs(v0d - v0o) <= 1e-12 * v0o
    clk = FrameClock(g["dx"], 1.0)
    opts = capi.step_opts(reinit_interval=2, mode=capi.PCG_FP64, tol=1e-12, max_iterations=20000)
    worst = 0.0
    for s in range(steps):
        dt = clk.next_dt(o.max_speed())
        r = o.step(dt, do_reinit=(s % 2 == 0), tol=1e-12)
        assert r["rc"] == 0
        d.step(dt, s, opts)
        drift_o = (o.volume() - v0o) / v0o
        drift_d = (d.volume() - v0d) / v0d
        worst = max(worst, abs(drift_d - drift_o))
    say(f"  final drift oracle {drift_o:.3e} gpu {drift_d:.3e}; worst difference {worst:.2e}")
    assert worst <= 1e-3


def _cube_volume(d, v):
    """GPU volume of one cube with circular-order corner values v (a 2x2x2
    grid of cells, unit spacing, corrected order)."""
    lex = np.empty(8)
    for c, (x, y, z) in enumerate(G.CIRC):
        lex[x + 2 * y + 4 * z] = v[c]
    d.set(capi.PHI_TMP, lex)
    return d.volume(corrected=True)


def test_ambiguous_cases_closed_form_on_gpu():
    say("volume: ambiguous corner configurations, closed forms, one cube per call")
    d = capi.Liquid(3, [2, 2, 2], 1.0, [0.0, 0.0, 0.0])
    rng = np.random.default_rng(21)
    worst = 0.0
    for m in range(256):
        if not G.corners_and_edges(m):
            continue
        for _ in range(3):
            alpha = rng.uniform(0.05, 5.0, 8)
            v = [-alpha[i] if m >> i & 1 else 1.0 for i in range(8)]
            worst = max(worst, abs(_cube_volume(d, v) - G.separated_volume(m, alpha)))
    for a, b in G.BODY_DIAGONALS:
        ba, bb = rng.uniform(0.05, 5.0, 2)
        v = [-1.0] * 8
        v[a], v[b] = ba, bb
        worst = max(worst, abs(_cube_volume(d, v) - G.body_diagonal_complement_volume(ba, bb)))
    say(f"  worst deviation {worst:.3e}")
    assert worst <= 1e-14
    # rotation invariance of the configurations that have it, on the GPU
    R = G.rotations()
    from test_volume_oracle import NOT_ROTATION_INVARIANT
    for m in range(256):
        if G.orbit_class(m) in NOT_ROTATION_INVARIANT:
            continue
        mag = rng.uniform(0.05, 1.0, 8)
        v = np.where([m >> i & 1 for i in range(8)], -mag, mag)
        want = _cube_volume(d, v)
        for p in R[1::5]:
            w = np.empty(8)
            w[p] = v
            assert abs(_cube_volume(d, w) - want) <= 1e-15, (m, p)
    d.close()
