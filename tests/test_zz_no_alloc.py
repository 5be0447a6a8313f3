"""The allocation-free volume entry point (garm_volume_device) against the
context entry point (garm_volume): same field, same result bit for bit, on
two calls in a row, the second on a smaller grid, so the shared scratch array
is reused with stale partial sums from the first call.

Device buffers come from the HIP runtime that libgarm_gpu.so already loaded
(through ctypes), not from torch: torch ships its own HIP runtime under the
same soname, and a process that loaded libgarm_gpu.so first cannot start
torch's device support."""
import ctypes

import numpy as np
import pytest

from helpers import say

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the process's HIP runtime."""

    def __init__(self):
        from garm_amd import capi
        capi.lib()                                   # loads the runtime libgarm_gpu.so links
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed with HIP error {rc}"

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        self._ok(self.h.hipMalloc(ctypes.byref(p), a.nbytes), "hipMalloc")
        self._ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def download(self, ptr, like):
        a = np.empty_like(like)
        self._ok(self.h.hipMemcpy(a.ctypes.data, ctypes.c_void_p(ptr), a.nbytes, 2), "hipMemcpy D2H")
        return a

    def free(self, ptr):
        self._ok(self.h.hipFree(ctypes.c_void_p(ptr)), "hipFree")


def _field(res, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(res))
    # smooth part (a sphere) plus noise: every inside/outside corner pattern occurs
    idx = np.indices(res[::-1]).reshape(len(res), -1)[::-1].T.astype(np.float64)
    c = (np.array(res) - 1) / 2.0
    r = np.sqrt(((idx - c) ** 2).sum(axis=1)) - 0.3 * min(res)
    return r + rng.normal(0.0, 1.5, n)


@pytest.mark.parametrize("dim,sizes", [(3, [(40, 36, 30), (18, 14, 12)]), (2, [(96, 80), (24, 20)])])
def test_volume_device_matches_context_volume(dim, sizes):
    from garm_amd import capi
    hip = _Hip()
    dx = 0.03125
    for call, res in enumerate(sizes):
        for corrected in (False, True):
            say(f"volume_device dim={dim} res={res} call={call} corrected={corrected}")
            phi = _field(res, seed=17 + call)
            d = capi.Liquid(dim, list(res), dx, [0.0] * dim)
            d.set(capi.PHI_TMP, phi)
            want = d.volume(corrected)
            d.close()
            nan = np.full(1, np.nan)
            phi_dev, out_dev = hip.upload(phi), hip.upload(nan)
            capi.volume_device(dim, res, dx, phi_dev, out_dev, corrected)   # null stream; hipMemcpy waits for it
            got = float(hip.download(out_dev, nan)[0])
            hip.free(phi_dev)
            hip.free(out_dev)
            assert got == want, (res, corrected, got, want)


def test_volume_device_rejects_oversize_and_bad_arguments():
    from garm_amd import capi
    hip = _Hip()
    phi, out = hip.upload(np.zeros(8)), hip.upload(np.zeros(1))
    try:
        with pytest.raises(capi.GarmError) as e:
            capi.volume_device(3, (5000, 2, 2), 0.1, phi, out)
        assert e.value.code == capi.GARM_EINVAL and "resolution" in str(e.value)
        with pytest.raises(capi.GarmError):
            capi.volume_device(4, (2, 2, 2), 0.1, phi, out)
        with pytest.raises(capi.GarmError):
            capi.volume_device(3, (2, 2, 2), 0.1, 0, out)
    finally:
        hip.free(phi)
        hip.free(out)
