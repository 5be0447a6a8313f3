"""Surface mesh on the GPU (garm_contour, garm_write_mesh) against the
oracle's Contourer<3>::contour restatement on the same phi_tmp: vertex
positions and triangle indices identical, normals within float rounding;
the liquid.mesh file read back equal to the in-memory mesh, and the OBJ
file equal, character for character, to the oracle's FP64 positions printed
as SurfaceMesh::writeOBJ prints them.  Fields: a
random one (every cube type occurs), a sphere, and a dam break's phi_tmp
after a few steps at 64^3; in 2D the filled contour on a random field and a
dam break.
"""
import os

import numpy as np
import pytest

import oracle as O
from garm_amd import capi, scenes
from garm_amd.driver import FrameClock
from helpers import make_gpu, say

pytestmark = pytest.mark.gpu


def _node_origin(g):
    return [g["origin"][a] + g["dx"] / 2 for a in range(3)]


def _compare(g, d, field):
    pos, nrm, tri = d.contour()
    opos, onrm, otri = O.contour3(field.reshape(g["res"][::-1]), _node_origin(g), g["dx"])
    say(f"  {len(pos)} vertices, {len(tri)} triangles")
    assert pos.shape == opos.shape and tri.shape == otri.shape
    assert np.array_equal(pos, opos)
    bad = np.nonzero((tri != otri).any(axis=1))[0]
    if len(bad):
        say(f"  {len(bad)} triangles differ, first {bad[:4]}: gpu {tri[bad[:4]].tolist()} oracle {otri[bad[:4]].tolist()}")
    assert len(bad) == 0
    assert np.max(np.abs(nrm - onrm)) <= 2e-7
    return pos, nrm, tri


def test_random_field_every_cube_type():
    g = scenes.grid(24, 3)
    d = make_gpu(g)
    rng = np.random.default_rng(11)
    f = rng.choice([-1.0, 1.0], size=int(np.prod(g["res"]))) * rng.uniform(0.2, 2.0, size=int(np.prod(g["res"])))
    d.set(capi.PHI_TMP, f)
    say("random field 24^3")
    _compare(g, d, f)


def test_sphere_and_file(tmp_path):
    g = scenes.grid(48, 3)
    d = make_gpu(g)
    c = np.array([0.45, 0.5, 0.55])
    f = np.linalg.norm(scenes.cell_positions(g) - c, axis=1) - 0.3
    d.set(capi.PHI_TMP, f)
    say("sphere 48^3")
    pos, nrm, tri = _compare(g, d, f)
    path, obj = os.path.join(tmp_path, "liquid.mesh"), os.path.join(tmp_path, "0.obj")
    d.write_mesh(path, obj)
    fp, fn, ft = capi.read_mesh(path)
    assert np.array_equal(fp, pos) and np.array_equal(fn, nrm) and np.array_equal(ft, tri)
    assert os.path.getsize(path) == 4 + 24 * len(pos) + 4 + 12 * len(tri)
    # SurfaceMesh::writeOBJ: FP64 positions printed %.15e, 1-based faces, no normals
    p64, _, _ = O.contour3(f.reshape(g["res"][::-1]), _node_origin(g), g["dx"], fp64_positions=True)
    want = "".join("v %.15e %.15e %.15e\n" % tuple(p) for p in p64) + "".join("f %d %d %d\n" % tuple(t + 1) for t in tri)
    assert open(obj).read() == want


def test_dam_break_phi_tmp():
    g = scenes.grid(64, 3)
    d = make_gpu(g)
    scenes.dam_break(d, g)
    d.initialize(capi.step_opts(tol=1e-8, max_iterations=20000))
    clk = FrameClock(g["dx"], 1.0)
    opts = capi.step_opts(reinit_interval=2, tol=1e-6, max_iterations=20000)
    for s in range(6):
        d.step(clk.next_dt(d.max_speed()), s, opts)
    say("dam break 64^3 after 6 steps")
    _compare(g, d, d.get(capi.PHI_TMP))


@pytest.mark.parametrize("n", [24, 64])
def test_filled_contour_2d(n, tmp_path):
    """2D: the filled contour (Contourer<2, true>), from the host copy of
    phi_tmp, against the oracle's, positions and indices identical; the
    liquid.mesh (float2 positions, no normals) and OBJ files."""
    g = scenes.grid(n, 2)
    d = make_gpu(g)
    if n == 24:
        rng = np.random.default_rng(5)
        f = rng.choice([-1.0, 1.0], size=int(np.prod(g["res"]))) * rng.uniform(0.2, 2.0, size=int(np.prod(g["res"])))
        d.set(capi.PHI_TMP, f)
    else:
        scenes.dam_break(d, g)
        d.initialize(capi.step_opts(tol=1e-8, max_iterations=20000))
        clk = FrameClock(g["dx"], 1.0)
        opts = capi.step_opts(reinit_interval=2, tol=1e-6, max_iterations=20000)
        for s in range(10):
            d.step(clk.next_dt(d.max_speed()), s, opts)
        f = d.get(capi.PHI_TMP)
    say(f"filled contour 2d n={n}")
    pos, nrm, tri = d.contour()
    o = [g["origin"][a] + g["dx"] / 2 for a in range(2)]
    opos, otri = O.contour2(f.reshape(g["res"][::-1]), o, g["dx"])
    say(f"  {len(pos)} vertices, {len(tri)} triangles")
    assert nrm.shape == (0, 3)
    assert np.array_equal(pos, opos.astype(np.float32)) and np.array_equal(tri, otri)
    path, obj = os.path.join(tmp_path, "liquid.mesh"), os.path.join(tmp_path, "0.obj")
    d.write_mesh(path, obj)
    fp, _, ft = capi.read_mesh(path, dim=2)
    assert np.array_equal(fp, pos) and np.array_equal(ft, tri)
    want = "".join("v %.15e %.15e 0\n" % tuple(p) for p in opos) + "".join("f %d %d %d\n" % tuple(t + 1) for t in otri)
    assert open(obj).read() == want
