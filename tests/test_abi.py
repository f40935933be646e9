"""The C-ABI library loads and exports every entry point include/swept1d.h
declares; argument validation works without a GPU (no compute calls)."""
import ctypes as C
import re

import pytest


def _declared(header):
    text = open(header).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(sw1d_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(sw):
    names = _declared(sw.HEADER)
    assert len(names) >= 18
    L = sw.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version(sw):
    assert sw.lib().sw1d_abi_version() == 1


@pytest.mark.parametrize("field,value,code", [
    ("equation", 7, -1), ("precision", 3, -1), ("algorithm", 9, -1), ("nx", 4, -1), ("dt", 0.0, -1),
])
def test_plan_validation_without_gpu(sw, field, value, code):
    prob = sw.Problem(1, 0, 1, 0, 1024, 64, 0, 0.005, 0.0, 0.0)
    setattr(prob, field, value)
    h = C.c_void_p()
    rc = sw.lib().sw1d_plan_create(C.byref(h), C.byref(prob))
    assert rc == code
    assert not h.value
    assert sw.last_error()


def test_bc_not_defined_for_equation(sw):
    prob = sw.Problem(1, 0, 1, sw.BC_MIRROR, 1024, 64, 0, 0.005, 0.0, 0.0)
    h = C.c_void_p()
    assert sw.lib().sw1d_plan_create(C.byref(h), C.byref(prob)) == -1
    assert "boundary" in sw.last_error()


def test_euler_unstable_dt_rejected(sw):
    # dt/dx = 0.5 > 0.21 (Euler1D_SweptShared.cu:1168-1172)
    n = 1024
    prob = sw.Problem(2, 0, 0, 0, n, 64, 0, 0.5 / (n - 2), 0.0, 0.0)
    h = C.c_void_p()
    assert sw.lib().sw1d_plan_create(C.byref(h), C.byref(prob)) == -4
    assert "too high" in sw.last_error()


def test_null_arguments(sw):
    L = sw.lib()
    assert L.sw1d_advance(None, 1) == -1
    assert L.sw1d_get_state(None, None) == -1
    assert L.sw1d_plan_destroy(None) == 0


@pytest.mark.parametrize("n,seed", [(1000, 1), (1 << 16, 7)])
def test_initial_condition_equals_oracle(sw, orc, n, seed):
    """sw1d_initial_condition is host-only code (no GPU): its seeded fields are
    the oracle's bit for bit, rounded once to float in single precision."""
    import numpy as np

    want = orc.ic_ks_noisy(n, seed)
    assert np.array_equal(sw.ks_noisy_ic(n, seed), want)
    assert np.array_equal(sw.ks_noisy_ic(n, seed, precision=sw.F32), want.astype(np.float32))
    for L in (1.0, n * 0.001):
        want, _ = orc.ic_heat_fourier(n, seed, L=L)
        assert np.array_equal(sw.heat_fourier_ic(n, seed, L=L), want)


def test_initial_condition_validation(sw):
    import numpy as np

    out = np.empty(4)
    L = sw.lib()
    assert L.sw1d_initial_condition(9, 0, 4, 0.5, 0.0, 1, out.ctypes.data_as(C.c_void_p)) == -1
    assert L.sw1d_initial_condition(0, 0, 0, 0.5, 0.0, 1, out.ctypes.data_as(C.c_void_p)) == -1
    assert L.sw1d_initial_condition(0, 0, 4, 0.5, 0.0, 1, None) == -1
