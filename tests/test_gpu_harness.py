"""The reference's test procedure (SweptSource/Testing/testProcedure.py) restated
through the Python 3 harness and the drop-in programs: consistency of the
algorithms, accuracy against the exact solutions and the KS fixture, the
full-precision snapshot / resume, and the seeded workload options checked
bit for bit against the CPU oracle."""
import importlib
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

h = importlib.import_module("1dsweptcuda_amd.harness")
FIXTURE = os.path.join(GOLDEN, "KSDouble_Official.txt")


def _run(problem, args, tmp_path, options=()):
    print("run", problem, *args, *options, flush=True)
    codes = h.runCUDA(h.program(problem), *args[:6], varfile=str(tmp_path / args[6]),
                      options=options, quiet=True)
    assert codes == [0]
    return h.Solved(str(tmp_path / args[6]))


@pytest.mark.parametrize("problem,tf,dt", [("Heat", 0.5, 1e-5), ("KS", 2.0, 1e-3), ("Euler", 0.01, 1e-5)])
def test_consistency(problem, tf, dt, tmp_path):
    """testProcedure.py:106-122 at div=4096, tpb=128: swept and the alternative
    algorithm against classic at the final time.  The reference passes with
    RMSE 0.000e+00 (consistency.out:2-13); here too, exactly."""
    print("consistency", problem, flush=True)
    rep, ok = h.consistency_report(problem, tf, dt=dt, workdir=str(tmp_path))
    assert ok and rep and all(v == 0.0 for v in rep.values()), rep


def test_accuracy_heat_series(tmp_path):
    """testProcedure.py:241-283 (classic, divs 1024, tpb 64): RMSE against the
    series solution in the published band 1.84e-2 .. 1.865e-2 at t = 15..50."""
    s = _run("Heat", [1024, 64, 1e-3, 50, 15, 0, "heat.dat"], tmp_path).stripInitial()["Temperature"]
    assert sorted(s) == [15.0, 30.0, 45.0, 50.0]
    for t, v in s.items():
        assert 1.80e-2 < h.rmse(h.heat_exact(t, 1024), v) < 1.90e-2, t


def test_accuracy_euler_sod(tmp_path):
    """Against the exact Sod solution at the published bands (nX = 1024,
    t up to 0.22): rho <= 8.8e-3, p <= 4.5e-3, u <= 1.8e-2, e <= 5.7e-2."""
    rows = _run("Euler", [1024, 64, 1e-5, 0.22, 0.066, 0, "eu.dat"], tmp_path).stripInitial()
    bands = {"Density": 8.8e-3, "Pressure": 4.5e-3, "Velocity": 1.8e-2, "Energy": 5.7e-2}
    assert sorted(rows) == sorted(bands)
    for name, band in bands.items():
        for t, v in rows[name].items():
            assert h.rmse(h.euler_exact(t, 1024, name), v) < band, (name, t)


def test_accuracy_ks_fixture(tmp_path):
    """KS against the reference's own dt=1e-7 solution: O(dt^2), RMSE 3.1e-8
    at t=6 for dt=1e-3 (the measured value of the restated stencil)."""
    rows = _run("KS", [1024, 64, 1e-3, 20, 6, 0, "ks.dat"], tmp_path).stripInitial()["Velocity"]
    for t, bound in ((6.0, 4.0e-8), (12.0, 1.7e-8), (18.0, 5e-9), (20.0, 3.3e-9)):
        assert h.rmse(h.ks_exact(t, 1024, "Velocity", FIXTURE), rows[t]) < bound, t


def test_snapshot_resume_is_a_true_restart(tmp_path, orc, sw):
    """A run resumed from a snapshot continues the first one bit for bit: both
    snapshots equal the oracle after N and 2N steps at full precision (which
    the 10-digit .dat cannot show), and the printed times continue."""
    args = [1024, 64, 1e-3, 4, 4, 1]
    _run("KS", args + ["a.dat"], tmp_path, ["--snapshot=" + str(tmp_path / "a.snap")])
    b = _run("KS", args + ["b.dat"], tmp_path,
             ["--resume=" + str(tmp_path / "a.snap"), "--snapshot=" + str(tmp_path / "b.snap")])
    ha, ua = h.read_snapshot(str(tmp_path / "a.snap"))
    hb, ub = h.read_snapshot(str(tmp_path / "b.snap"))
    steps, times = sw.schedule(3, 1e-3, 4, 4, 64)  # KS swept bookkeeping
    assert ha["time"] == times[-1] and hb["time"] == 2 * times[-1]
    assert b.tFinal[0] == np.around(times[-1], 7)
    c = orc.ks_consts(1e-3)
    ua_want = orc.ks_run(orc.ks_ic_ref(1024), steps[-1], c)
    assert np.array_equal(ua, ua_want)
    assert np.array_equal(ub, orc.ks_run(ua_want, steps[-1], c))


def test_seeded_workload_options_match_oracle(tmp_path, orc, sw):
    """--ic=noisy (KS) and --ic=fourier --bc=periodic (heat) draw the same
    splitmix64 values as the oracle's generators and advance bit for bit."""
    _run("KS", [2048, 128, 5e-3, 1, 1, 1, "k.dat"], tmp_path,
         ["--ic=noisy", "--seed=3", "--snapshot=" + str(tmp_path / "k.snap")])
    hk, uk = h.read_snapshot(str(tmp_path / "k.snap"))
    ks_steps = sw.schedule(3, 5e-3, 1, 1, 128)[0][-1]  # KS swept bookkeeping
    want = orc.ks_run(orc.ic_ks_noisy(2048, 3), ks_steps, orc.ks_consts(5e-3))
    assert np.array_equal(uk, want)

    _run("Heat", [1024, 64, 1e-3, 2, 2, 1, "p.dat"], tmp_path,
         ["--ic=fourier", "--seed=5", "--bc=periodic", "--snapshot=" + str(tmp_path / "p.snap")])
    hp, up = h.read_snapshot(str(tmp_path / "p.snap"))
    steps = sw.schedule(1, 1e-3, 2, 2, 64)[0][-1]  # heat swept bookkeeping
    fo, fo2 = orc.heat_consts(1e-3)
    u0, _ = orc.ic_heat_fourier(1024, 5, L=1024 * 0.001)
    assert np.array_equal(up, orc.heat_run(u0, steps, fo, fo2, 1))


def test_sod_to_t02_at_2p14(sw):
    """The Euler configuration at its smallest size: Sod, nX = 2^14,
    dt/dx = 0.1, run to t = 0.2 (32764 steps; the exact solution is taken at
    the quantised time).  Swept and classic give the same bits, and every
    variable is closer to the exact Riemann solution than the published
    nX = 1024 bands allow, by at least half (first-order shocks converge)."""
    n = 1 << 14
    dx = 1.0 / (n - 2)
    dt = 0.1 * dx
    steps = int(np.ceil(0.2 / dt))
    q0 = np.zeros((n, 3))
    q0[: n // 2] = (1.0, 0.0, 1.0 / 0.4)
    q0[n // 2:] = (0.125, 0.0, 0.1 / 0.4)
    out = {}
    for alg, W in ((1, 512), (0, 256)):
        print(f"Sod nX=2^14 alg={alg} steps={steps}", flush=True)
        with sw.Plan(2, n, dt, algorithm=alg, tpb=W, node_width=W if alg else 0) as p:
            p.set_state(q0.ravel())
            p.advance(steps)
            out[alg] = p.get_state().reshape(n, 3)
    assert np.array_equal(out[0], out[1])
    q = out[1][1:-1]
    rho, m, E = q[:, 0], q[:, 1], q[:, 2]
    u = m / rho
    got = {"Density": rho, "Velocity": u, "Energy": E / rho - 0.5 * u * u, "Pressure": 0.4 * (E - 0.5 * m * m / rho)}
    bands = {"Density": 8.8e-3, "Pressure": 4.5e-3, "Velocity": 1.8e-2, "Energy": 5.7e-2}
    for name, band in bands.items():
        err = h.rmse(h.euler_exact(steps * dt, n, name), got[name])
        print(f"  {name} RMSE {err:.3e}", flush=True)
        assert err < 0.5 * band, (name, err)


def test_timing_sweep_files(tmp_path):
    """The timing sweep (runTest.sh:8-38) on a small grid: one file per
    problem, precision and scheme with the reference's header, one row per
    (nX, tpb), read back by gatherTest and QualityRuns."""
    print("timing sweep KS, nX 2048 and 4096, tpb 64 and 128, 200 steps", flush=True)
    files = h.timing_sweep(str(tmp_path), nts=200, divisions=[2048, 4096], threadsPerBlock=[64, 128],
                           problems=("KS",), progress=None)
    assert [os.path.basename(f) for f in files] == ["KS_Double_%s_Timing.txt" % s for s in h.SWEEP_SCHEMES]
    for f in files:
        lines = open(f).read().splitlines()
        assert lines[0] == h.TIMING_HEADER.strip("\n") and len(lines) == 5
    frame = h.gatherTest(str(tmp_path))
    assert frame.shape == (6, 2) and (frame.values > 0).all()
    best = h.QualityRuns(frame).bestRun
    assert len(best) == 6
    assert not os.path.exists(tmp_path / "temp.dat")
