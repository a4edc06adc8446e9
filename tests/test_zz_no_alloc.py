"""ik_mlp_forward_noalloc (no hipMalloc: hidden activations in the library's fixed per-device scratch array)
against ik_mlp_forward on the same inputs, through the C ABI: two calls in a row, the second with a smaller batch,
results equal bit for bit; batches beyond IK_MLP_NOALLOC_MAX_ROWS are rejected with IK_EINVAL."""
import ctypes as C
import os

import numpy as np
import pytest

from ikarus_amd import _capi as capi
from ikarus_amd import learned

pytestmark = pytest.mark.gpu

MAX_ROWS = 8192
MLP_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ikarus_amd", "data",
                        "predictor_mlp_32cubed.f32")


class _Hip:
    """Device buffers for the device-pointer entry point (the runtime's own allocator, outside the library)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.lib.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.lib.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2) == 0


def _features(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, (n, learned.N_IN)).astype(np.float32)


def test_noalloc_forward_matches_forward_two_calls():
    print("case: mlp forward noalloc", flush=True)
    hip = _Hip()
    m = learned.MLP(seed=3)
    m.load(np.fromfile(MLP_PATH, dtype=np.float32))
    xd = hip.malloc(MAX_ROWS * learned.N_IN * 4)
    yd = hip.malloc(MAX_ROWS * learned.N_OUT * 4)
    try:
        for n, seed in ((MAX_ROWS, 1), (1000, 2)):   # the second call is smaller: stale scratch rows are not read
            x = _features(n, seed)
            ref = m.forward(x)
            hip.h2d(xd, x)
            assert hip.lib.hipMemset(yd, 0xFF, MAX_ROWS * learned.N_OUT * 4) == 0
            m._check(capi.lib().ik_mlp_forward_noalloc(m.h, xd, n, yd))
            y = np.zeros((n, learned.N_OUT), np.float32)
            hip.d2h(y, yd)
            assert np.array_equal(y, ref), (n, np.abs(y - ref).max())
        for bad in (0, MAX_ROWS + 1):
            rc = capi.lib().ik_mlp_forward_noalloc(m.h, xd, bad, yd)
            assert rc == capi.IK_EINVAL
        assert b"IK_MLP_NOALLOC_MAX_ROWS" in capi.lib().ik_mlp_last_error(m.h)
        rc = capi.lib().ik_mlp_forward_noalloc(m.h, None, 10, yd)
        assert rc == capi.IK_EINVAL
    finally:
        hip.lib.hipFree(xd)
        hip.lib.hipFree(yd)
        m.close()
