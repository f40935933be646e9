"""The swept decomposition (tiles, offsets, active sets) reproduces the classic
scheme bitwise — the reference's own consistency invariant
(ResultPlots/ExactTesting/consistency.out: swept vs classic RMSE 0.000e+00) —
checked on the CPU with the emulator of the kernels' tile schedule."""
import numpy as np
import pytest

from swept_emulator import Emu


def _ks_ic(orc, n, dtype):
    return orc.ks_ic_ref(n, dtype) + np.asarray(orc.ic_ks_noisy(n, 7) - orc.ks_ic_ref(n), dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("W,nsteps", [(64, 32 * 5), (128, 64 * 3 + 17), (256, 128)])
@pytest.mark.parametrize("bc", ["mirror", "periodic"])
def test_heat_swept_equals_classic(orc, dtype, W, nsteps, bc):
    n = 1024
    dt = 0.001 if bc == "mirror" else 1.0 / 1024 ** 2 * 0.25 / 8.418e-5
    e = Emu("heat", n, dt, dtype, bc=bc, W=W)
    if bc == "mirror":
        u0 = orc.heat_ic_ref(n, dtype)
    else:
        u0 = orc.ic_heat_fourier(n, 3)[0].astype(dtype)
    sw = e.run(u0, nsteps)
    cl = orc.heat_run(u0, nsteps, e.fo, e.fo2, 0 if bc == "mirror" else 1)
    assert np.array_equal(sw, cl)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("W,nsteps", [(128, 16 * 5), (256, 32 * 3 + 5), (512, 64)])
def test_ks_swept_equals_classic(orc, dtype, W, nsteps):
    n = 2048
    e = Emu("ks", n, 0.005, dtype, W=W)
    u0 = _ks_ic(orc, n, dtype)
    sw = e.run(u0, nsteps)
    cl = orc.ks_run(u0, nsteps, e.coef)
    assert np.array_equal(sw, cl)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("W,nsteps", [(128, 16 * 7), (256, 32 * 2 + 3)])
def test_euler_swept_equals_classic(orc, dtype, W, nsteps):
    n = 1024
    dx = 1.0 / (n - 2)
    e = Emu("euler", n, 0.1 * dx, dtype, W=W)
    q0 = orc.euler_ic_ref(n, dtype)
    sw = e.run(q0, nsteps)
    cl = orc.euler_run(q0, nsteps, e.k, e.bd)
    assert np.array_equal(sw, cl)


def test_emulator_detects_a_wrong_active_set(orc, monkeypatch):
    """A plausible geometry mistake (one point too many in the down phase)
    must break the equality, so the tests above have teeth."""
    import swept_emulator as se

    n, W = 1024, 128
    e = Emu("ks", n, 0.005, np.float64, W=W)
    u0 = _ks_ic(orc, n, np.float64)
    good = orc.ks_run(u0, 32, e.coef)
    real_mask = se._mask
    monkeypatch.setattr(se, "_mask", lambda n_, W_, sh, thr: real_mask(n_, W_, sh, thr + 2))
    bad = e.run(u0, 32)
    assert not np.array_equal(bad, good)
