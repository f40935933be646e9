"""sw1d_advance_noalloc, the allocation-free advance, through the C ABI: each
call gives the bits of the plan path (sw1d_plan_create + set_state + advance +
get_state) for the same problem, on two calls in a row where the second has
the smaller input, so the workspace still holds the first call's contents.
Free device memory (hipMemGetInfo) does not fall across the second call.  Runs
after the other GPU tests and alone."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    return C.CDLL("libamdhip64.so")


def _free_bytes(hip):
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _initial(orc, sw, eq, prec, n, bc):
    dtype = np.float64 if prec == sw.F64 else np.float32
    if eq == sw.KS:
        return orc.ic_ks_noisy(n, 5).astype(dtype)
    if eq == sw.EULER:
        return orc.euler_ic_ref(n, dtype)
    if bc == sw.BC_PERIODIC:
        return orc.ic_heat_fourier(n, 3)[0].astype(dtype)
    return orc.heat_ic_ref(n, dtype)


def _dt(sw, eq, n):
    return {sw.KS: 0.005, sw.HEAT: 1e-3, sw.EULER: 0.1 / (n - 2)}[eq]


# (equation, precision, algorithm, first nx, second nx, node width, tile depth, bc, steps)
CASES = [
    ("ks-f64-swept", 1, 0, 1, 1 << 16, 1 << 14, 256, 0, 0, 1000),
    ("ks-f32-register", 1, 1, 2, 1 << 15, 4096, 0, 0, 0, 700),
    ("heat-f64-swept-periodic", 0, 0, 1, 1 << 14, 2048, 512, 0, 2, 1500),
    ("heat-f64-classic-mirror", 0, 0, 0, 4096, 1024, 0, 0, 0, 301),
    ("euler-f64-swept", 2, 0, 1, 1 << 14, 4096, 256, 0, 0, 900),
    ("euler-f32-classic", 2, 1, 0, 8192, 2048, 0, 0, 0, 333),
    ("ks-f64-trapezoid-ragged", 1, 0, 3, 1 << 15, 5000, 512, 2, 0, 777),
]


def _noalloc(sw, hip, prob, u0, steps):
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), C.c_size_t(u0.nbytes)) == 0
    try:
        assert hip.hipMemcpy(d, u0.ctypes.data_as(C.c_void_p), C.c_size_t(u0.nbytes), H2D) == 0
        assert hip.hipDeviceSynchronize() == 0
        free0 = _free_bytes(hip)
        rc = sw.lib().sw1d_advance_noalloc(C.byref(prob), d, steps, None)
        assert rc == 0, sw.last_error()
        assert hip.hipDeviceSynchronize() == 0
        free1 = _free_bytes(hip)
        out = np.empty_like(u0)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d, C.c_size_t(u0.nbytes), D2H) == 0
    finally:
        hip.hipFree(d)
    return out, free0, free1


def _declare(sw):
    L = sw.lib()
    L.sw1d_advance_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    return L


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_plan_path(sw, orc, case):
    name, eq, prec, alg, n1, n2, W, depth, bc, steps = case
    hip = _hip()
    _declare(sw)
    for call, n in enumerate((n1, n2)):
        print(f"noalloc {name} call {call} nx={n} steps={steps}", flush=True)
        dt = _dt(sw, eq, n)
        u0 = _initial(orc, sw, eq, prec, n, bc)
        with sw.Plan(eq, n, dt, precision=prec, algorithm=alg, tpb=W or 64, node_width=W, bc=bc,
                     tile_depth=depth) as p:
            p.set_state(u0)
            p.advance(steps)
            want = p.get_state()
            prob = p.prob
        got, free0, free1 = _noalloc(sw, hip, prob, u0, steps)
        assert np.array_equal(got, want), f"{int(np.sum(got != want))} values differ"
        assert not np.array_equal(got, u0)
        if call == 1:  # the first call on a device may load the library's device code
            # free memory can grow (the runtime releasing the plan just destroyed), never shrink
            assert free1 >= free0, f"device memory fell by {free0 - free1} bytes"


def test_noalloc_zero_steps_and_errors(sw, orc):
    hip = _hip()
    L = _declare(sw)
    n = 4096
    u0 = orc.ic_ks_noisy(n, 1)
    prob = sw.Problem(sw.KS, sw.F64, sw.SWEPT, 0, n, 256, 256, 0.005, 0.0, 0.0, 0, 0)
    got, _, _ = _noalloc(sw, hip, prob, u0, 0)
    assert np.array_equal(got, u0)
    d = C.c_void_p(1)  # never dereferenced: every case below is rejected before any device work
    big = sw.Problem(sw.KS, sw.F64, sw.SWEPT, 0, 1 << 22, 1024, 1024, 0.005, 0.0, 0.0, 0, 0)
    assert L.sw1d_advance_noalloc(C.byref(big), d, 10, None) == -1
    assert "workspace" in sw.last_error()
    hybrid = sw.Problem(sw.HEAT, sw.F64, sw.ALT, 0, 4096, 256, 256, 1e-3, 0.0, 0.0, 0, 0)
    assert L.sw1d_advance_noalloc(C.byref(hybrid), d, 10, None) == -5
    assert L.sw1d_advance_noalloc(C.byref(prob), None, 10, None) == -1
    assert L.sw1d_advance_noalloc(C.byref(prob), d, -1, None) == -1
