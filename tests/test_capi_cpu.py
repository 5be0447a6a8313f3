"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, and refuses to run without a gfx950 device."""
import ctypes

import pytest

from garm_amd import capi


def test_header_symbols_exported():
    L = capi.lib()
    names = capi.declared_symbols()
    assert "garm_create" in names and "garm_step" in names and "garm_project" in names
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/garm_gpu.h but not exported"


def test_create_rejects_bad_arguments():
    L = capi.lib()
    h = ctypes.c_void_p()
    res = (ctypes.c_int * 3)(8, 8, 8)
    org = (ctypes.c_double * 3)(0, 0, 0)
    assert L.garm_create(ctypes.byref(h), 4, res, 0.1, org, 2) == capi.GARM_EINVAL
    assert L.garm_create(ctypes.byref(h), 3, res, -1.0, org, 2) == capi.GARM_EINVAL
    assert not h.value


def test_no_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.GarmError) as e:
        capi.Liquid(3, [8, 8, 8], 0.25, [-0.5] * 3)
    assert e.value.code in (capi.GARM_ENODEV,)
