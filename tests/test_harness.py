"""The Python 3 analysis harness (1dsweptcuda_amd/harness.py) on files alone,
and the programs' option errors that stop before any device call.  CPU only."""
import importlib
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN

h = importlib.import_module("1dsweptcuda_amd.harness")


def test_solved_reads_reference_fixture():
    s = h.Solved(os.path.join(GOLDEN, "KSDouble_Official.txt"))
    assert s.numpts == 1024 and s.vals.shape == (5, 1024)
    assert list(s.tFinal) == [0, 6, 12, 18, 20]
    assert s.xGrid[-1] == 512.0 and set(s.varNames) == {"Velocity"}
    # row 0 is the reference IC 2cos(19 pi x/128), x = k dx, printed with 10 digits
    x = np.arange(1024) * 0.5
    assert np.max(np.abs(s.vals[0] - 2 * np.cos(19 * np.pi * x / 128))) < 1e-9
    rows = s.stripInitial()
    assert sorted(rows["Velocity"]) == [6, 12, 18, 20]
    assert np.array_equal(h.ks_exact(12.0, 1024, "Velocity", os.path.join(GOLDEN, "KSDouble_Official.txt")),
                          s.vals[2])


def _timing_file(path, rows):
    h.write_timing_header(path)
    with open(path, "a") as f:
        for r in rows:
            f.write("%d\t%d\t%g\n" % r)


def test_perform_pivots_and_quality_runs(tmp_path):
    p = tmp_path / "KS_Double_Swept_1.txt"
    _timing_file(p, [(1024, 64, 3.0), (1024, 128, 2.0), (2048, 64, 5.0), (2048, 128, 6.0)])
    m = h.Perform(str(p)).dataMatrix
    assert list(m.index) == [1024, 2048] and list(m.columns) == [64, 128]
    assert m.loc[1024, 128] == 2.0 and m.loc[2048, 64] == 5.0
    q = h.QualityRuns(m)
    assert list(q.bestRun.iloc[:, 0]) == [2.0, 5.0]
    assert dict(q.bestLaunch) == {64: 1, 128: 1}
    _timing_file(tmp_path / "Heat_Double_Classic_1.txt", [(1024, 64, 9.0)])
    g = h.gatherTest(str(tmp_path))
    assert g.index.names == ["Problem", "Precision", "Algorithm", "Num Spatial Points"]
    assert g.loc[("KS", "Double", "Swept", 2048), 128] == 6.0
    assert g.loc[("Heat", "Double", "Classic", 1024), 64] == 9.0


def test_read_snapshot_layout(tmp_path):
    vals = np.arange(12, dtype=np.float64).reshape(4, 3)
    p = tmp_path / "s.snap"
    with open(p, "wb") as f:
        f.write(struct.pack("<8siiqqdd", b"SW1DSNP1", 2, 0, 4, 3, 0.25, 1e-5))
        f.write(vals.tobytes())
    head, state = h.read_snapshot(str(p))
    assert head == {"equation": 2, "precision": 0, "nx": 4, "components": 3, "time": 0.25, "dt": 1e-5}
    assert np.array_equal(state, vals)
    with open(p, "r+b") as f:
        f.write(b"NOTASNAP")
    with pytest.raises(ValueError):
        h.read_snapshot(str(p))


def test_heat_series_starts_at_the_parabola():
    """At t = 0 the series of testProcedure.py:63-82 sums to the program's
    initial parabola 100 x (1 - x/L) on its own grid L = divs*dx (Q8)."""
    divs = 1024
    L = divs * 0.001
    x = np.linspace(0, L, divs)
    got = h.heat_exact(0.0, divs)
    assert np.max(np.abs(got - 100 * x * (1 - x / L))) < 1e-2
    late = h.heat_exact(1e9, divs)  # insulated ends: the mean, 100 L / 6, remains
    assert np.allclose(late, 100 * L / 6, atol=1e-9)


def test_sod_exact_on_program_cells():
    _, (p, u, rl, rr) = h.sod_exact(np.array([0.5]), 0.1)
    assert abs(p - 0.30313) < 1e-5 and abs(u - 0.92745) < 1e-5  # Toro, table 4.2
    rho = h.euler_exact(0.2, 1024, "Density")
    assert rho.size == 1022 and rho[0] == 1.0 and rho[-1] == 0.125
    assert np.all(np.diff(rho) <= 1e-15)  # non-increasing left to right for Sod
    vel = h.euler_exact(0.2, 1024, "Velocity")
    assert vel[0] == 0.0 and abs(vel.max() - u) < 1e-12


def _exe(name):
    p = h.program(name)
    if not os.path.exists(p):
        pytest.skip("programs not built")
    return p


@pytest.mark.parametrize("prog,args,msg", [
    ("KS", ["--bogus=1"], "INVALID OPTION: --bogus=1"),
    ("KS", ["--ic=fourier"], "INVALID OPTION: --ic=fourier"),
    ("Euler", ["--ic=noisy"], "INVALID OPTION: --ic=noisy"),
    ("Heat", ["--bc=dirichlet"], "INVALID OPTION: --bc=dirichlet"),
    ("Heat", ["--resume=/nonexistent/x.snap"], "does not hold this program's"),
])
def test_program_option_errors(tmp_path, prog, args, msg):
    base = {"KS": ["1024", "64", "1e-3", "1", "1", "1"], "Heat": ["1024", "64", "1e-3", "1", "1", "0"],
            "Euler": ["1024", "64", "1e-5", "0.01", "0.01", "0"]}[prog]
    r = subprocess.run([_exe(prog)] + base + [str(tmp_path / "o.dat")] + args, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 255 and msg in r.stdout, r.stdout + r.stderr
