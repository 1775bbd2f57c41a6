"""The allocation-free emulator forward (cs_mlp_forward_noalloc) against the
allocating one (cs_mlp_forward): bit-identical outputs on two calls in a row,
the second with a smaller batch (the shared scratch array keeps the first
call's contents, which must not leak into the second), and a batch beyond the
scratch array's size refused on the host before any launch.

Device buffers for x and y come from the HIP runtime the library itself links
(libamdhip64), called through ctypes; the test allocates them, the entry point
under test does not.
"""
import ctypes as C
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, a):
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def test_noalloc_forward_equals_forward():
    emod = importlib.import_module("covid-sim_amd.emulator")
    engine = importlib.import_module("covid-sim_amd.engine")
    n_in, n_h, n_out = 10, 512, 730
    emu = emod.Emulator(n_in, n_h, n_out, seed=5)
    rng = np.random.default_rng(11)
    emu.set_params((rng.standard_normal(emu.n_params) * 0.05).astype(np.float32))
    hip = _hip()
    big = 6000
    xd = DevBuf(hip, big * n_in * 4)
    yd = DevBuf(hip, big * n_out * 4)
    try:
        for batch in (big, 700):  # the second call is smaller
            x = rng.standard_normal((batch, n_in)).astype(np.float32)
            want = emu.forward(x)
            xd.put(x)
            yd.put(np.full((batch, n_out), np.nan, np.float32))  # no stale output can pass
            emu.forward_device(xd.p.value, batch, yd.p.value)
            got = yd.get(np.zeros((batch, n_out), np.float32))
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), batch
            assert np.isfinite(got).all()
        # beyond the scratch array: refused on the host, nothing launched
        with pytest.raises(engine.CSError, match="scratch"):
            emu.forward_device(xd.p.value, 20000, yd.p.value)
        # and the emulator still serves a valid batch afterwards
        x = rng.standard_normal((64, n_in)).astype(np.float32)
        xd.put(x)
        emu.forward_device(xd.p.value, 64, yd.p.value)
        got = yd.get(np.zeros((64, n_out), np.float32))
        assert np.array_equal(got.view(np.uint32), emu.forward(x).view(np.uint32))
    finally:
        xd.free()
        yd.free()
        emu.close()
