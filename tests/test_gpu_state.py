"""State transfer through the C ABI: host and device-pointer forms, both
layouts (Euler's interleaved (rho, rho*u, E) cells against the plan's
component arrays), FP64 and FP32, at a size where a per-element copy path
would be visible; the device buffer comes from hipMalloc."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _DeviceBuffer:
    """A hipMalloc allocation (libamdhip64 is already loaded by libswept1d)."""

    H2D, D2H = 1, 2

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.nbytes = nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0

    def read(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes),
                                  self.D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.mark.parametrize("eq", [1, 2])
@pytest.mark.parametrize("prec", [0, 1])
def test_state_round_trips(sw, orc, eq, prec):
    n = 1 << 18
    dtype = np.float64 if prec == 0 else np.float32
    dt = 0.005 if eq == 1 else 0.1 / (n - 2)
    u0 = orc.ic_ks_noisy(n, 4).astype(dtype) if eq == 1 else orc.euler_ic_ref(n, dtype)
    print(f"state round trip eq={eq} prec={prec} n={n}", flush=True)
    dev = _DeviceBuffer(u0.nbytes)
    try:
        with sw.Plan(eq, n, dt, precision=prec, algorithm=1, tpb=256) as p:
            p.set_state(u0)
            assert np.array_equal(p.get_state(), u0)
            p.advance(64)
            after = p.get_state()
            p.get_state_device(dev.ptr.value)
            p.synchronize()
        assert np.array_equal(dev.read(dtype, u0.size), after)
        with sw.Plan(eq, n, dt, precision=prec, algorithm=1, tpb=256) as q, \
                sw.Plan(eq, n, dt, precision=prec, algorithm=0) as r:
            q.set_state_device(dev.ptr.value)
            assert np.array_equal(q.get_state(), after)
            q.advance(64)
            r.set_state(after)
            r.advance(64)
            assert np.array_equal(q.get_state(), r.get_state())
    finally:
        dev.free()
