"""wx_sur_infer_noalloc (the surrogate inference for callers that cannot
allocate: device-resident model, inputs and outputs, the device model built
in a static scratch buffer) against wx_sur_infer, bit for bit, on two calls
in a row, the second with fewer samples and another model, so that a stale
model or stale outputs in the shared buffer would show."""
import ctypes

import numpy as np
import pytest

from walicxe2d_amd import surrogate as S
from walicxe2d_amd.capi import WxSurModel, load_library

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime, for the test's own
    device buffers (the entry point under test allocates nothing)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(int(nbytes), 1)) == 0
        self.bufs.append(p)
        return p

    def h2d(self, dst, src, nbytes):
        assert self.L.hipMemcpy(dst, src, int(nbytes), 1) == 0

    def d2h(self, dst, src, nbytes):
        assert self.L.hipMemcpy(dst, src, int(nbytes), 2) == 0   # synchronous: waits for the stream's kernels

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _model(seed):
    rng = np.random.default_rng(seed)
    m = S.SurrogateModel(seed=seed)
    m.theta = (m.theta + rng.normal(scale=0.05, size=m.theta.shape)).astype(np.float32)
    return m


def test_noalloc_matches_infer():
    L = load_library()
    H = _Hip()
    try:
        rng = np.random.default_rng(11)
        for call, (n, seed) in enumerate([(5000, 21), (1234, 22)]):
            m = _model(seed)
            x = rng.normal(size=(n, 4)).astype(np.float32)
            want = S.infer(m, x)                       # existing entry point
            cm = m.to_c()
            dm = H.alloc(ctypes.sizeof(WxSurModel))
            dx = H.alloc(x.nbytes)
            dy = H.alloc(n * 2 * 4)
            H.h2d(dm, ctypes.addressof(cm), ctypes.sizeof(WxSurModel))
            H.h2d(dx, x.ctypes.data, x.nbytes)
            print(f"call {call}: n={n} model seed {seed}", flush=True)
            rc = L.wx_sur_infer_noalloc(dm, dx, n, dy, None)
            assert rc == 0, rc
            H.sync()
            got = np.empty((n, 2), np.float32)
            H.d2h(got.ctypes.data, dy, got.nbytes)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (call, np.abs(got - want).max())
        # arguments the entry point refuses
        assert L.wx_sur_infer_noalloc(None, dx, 10, dy, None) != 0
        assert L.wx_sur_infer_noalloc(dm, dx, -1, dy, None) != 0
    finally:
        H.sync()
        H.free()
