"""Checkpoint and resume through the reference's .sav files
(LevelSetLiquid::saveFrame / loadFrame, LevelSetLiquid.cpp:425-526).

* Resume equivalence: N + M steps in one context equal N steps, save, a fresh
  context that loads the files, M steps -- bit for bit, for the plain FP64
  path and for the FP32 + multigrid path.
* Layout: the files hold the reference's raw arrays (IO::writeArray), so
  after N lock-step steps they equal the oracle's fields in its own
  array-of-structures layout; an oracle loaded from the GPU's files and run
  M more steps matches the GPU run within 1e-10 (files cross programs).
"""
import os

import numpy as np
import pytest

from garm_amd import capi, scenes
from garm_amd.driver import FrameClock
from helpers import copy_state, make_gpu, make_oracle, rel_linf, say

pytestmark = pytest.mark.gpu

SAV = [("velocity.sav", capi.VELOCITY, np.float64), ("refMap.sav", capi.REFMAP, np.float64),
       ("refLevelSet.sav", capi.PHI_REF, np.float64), ("advRegion.sav", capi.ADV_REGION, np.uint8),
       ("refMapGrad.sav", capi.REFMAP_GRAD, np.float64), ("refLevelSetGrad.sav", capi.PHI_REF_GRAD, np.float64)]
STATE = [capi.PHI, capi.PHI_GRAD, capi.VELOCITY, capi.PRESSURE, capi.REFMAP, capi.REFMAP_GRAD, capi.PHI_TMP,
         capi.PHI_ORI, capi.PHI_REF, capi.PHI_REF_GRAD, capi.ADV_REGION]


@pytest.mark.parametrize("dim,n,mode", [(2, 48, "fp64"), (3, 32, "fp64"), (3, 32, "fp32-multigrid")])
def test_resume_is_bit_identical(dim, n, mode, tmp_path):
    g = scenes.grid(n, dim)
    if mode == "fp64":
        opts = capi.step_opts(reinit_interval=3, tol=1e-10, max_iterations=50000)
    else:
        opts = capi.step_opts(reinit_interval=3, precision=capi.FP32, mode=capi.PCG_MULTIGRID, tol=1e-8)
    N, M = 7, 6
    a = make_gpu(g)
    scenes.dam_break(a, g)
    a.initialize(opts)
    clk = FrameClock(g["dx"], 1.0)
    dts = []
    for s in range(N):
        dts.append(clk.next_dt(a.max_speed()))
        a.step(dts[-1], s, opts)
    a.save_frame(str(tmp_path))
    for s in range(N, N + M):
        dts.append(clk.next_dt(a.max_speed()))
        a.step(dts[-1], s, opts)
    b = make_gpu(g)
    b.load_frame(str(tmp_path))
    for s in range(N, N + M):
        say(f"resume {dim}d n={n} {mode} step {s}")
        b.step(dts[s], s, opts)
    for f in STATE:
        x, y = a.get(f), b.get(f)
        assert np.array_equal(x, y), f"{capi.FIELD_NAMES[f]} differs after resume ({np.count_nonzero(x != y)} entries)"


def test_sav_files_match_the_oracle_and_resume_across_programs(tmp_path):
    g = scenes.grid(32, 3)
    o, d = make_oracle(g), make_gpu(g)
    for s in (o, d):
        scenes.dam_break(s, g)
    assert o.initialize(tol=1e-12, max_it=50000) == 0
    opts = capi.step_opts(reinit_interval=2, tol=1e-12, max_iterations=50000)
    clk = FrameClock(g["dx"], 1.0)
    for s in range(5):
        copy_state(o, d)
        dt = clk.next_dt(o.max_speed())
        o.step(dt, do_reinit=(s % 2 == 0), tol=1e-12, max_it=50000)
        d.step(dt, s, opts)
    d.save_frame(str(tmp_path))
    for name, f, dtype in SAV:
        raw = np.fromfile(os.path.join(tmp_path, name), dtype=dtype)
        assert raw.size == o.count(f), name
        ref = o.get(f)
        if dtype == np.uint8:
            assert np.array_equal(raw, ref), name
        else:
            assert rel_linf(raw, ref) <= 1e-10, name
    # the oracle resumes from the GPU's files; both run 4 more steps
    o2 = make_oracle(g)
    for name, f, dtype in SAV:
        o2.set(f, np.fromfile(os.path.join(tmp_path, name), dtype=dtype))
    for s in range(5, 9):
        dt = clk.next_dt(d.max_speed())
        say(f"cross-program resume step {s}")
        assert o2.step(dt, do_reinit=(s % 2 == 0), tol=1e-12, max_it=50000)["rc"] == 0
        d.step(dt, s, opts)
    for f in (capi.PHI, capi.VELOCITY, capi.PRESSURE, capi.REFMAP, capi.REFMAP_GRAD):
        e = rel_linf(d.get(f), o2.get(f))
        assert e <= 1e-9, f"{capi.FIELD_NAMES[f]} {e:.2e}"


def test_load_rejects_a_file_of_the_wrong_size(tmp_path):
    g = scenes.grid(16, 2)
    d = make_gpu(g)
    d.save_frame(str(tmp_path))
    with open(os.path.join(tmp_path, "refMap.sav"), "ab") as fh:
        fh.write(bytes(8))
    with pytest.raises(capi.GarmError) as e:
        d.load_frame(str(tmp_path))
    assert e.value.code == capi.GARM_EINVAL and "refMap.sav" in str(e.value)
    with pytest.raises(capi.GarmError):
        d.load_frame(str(tmp_path / "missing"))
