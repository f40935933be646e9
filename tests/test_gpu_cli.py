"""The drop-in programs bin/{Heat,KS,Euler}{Single,Double}Out: argv, .dat and
.txt formats, exit codes, and values — every printed number equal, as text,
to the oracle's value at the step count the reference's bookkeeping reaches
(README.md:56-68, Heat:697-836, KS:866-1029, Euler:1112-1287)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _run(sw, eq, prec, args, cwd):
    exe = sw.program(eq, prec)
    print("run", os.path.basename(exe), " ".join(map(str, args)), flush=True)
    return subprocess.run([exe] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True,
                          timeout=120)


def _rows(path):
    lines = open(path).read().split("\n")
    header = lines[0].split(" ")
    rows = [l for l in lines[1:] if l.strip()]
    return header, rows


def _fmt(v):
    return "%.10g" % float(v)


def test_ks_reproduces_reference_fixture(sw, tmp_path):
    """KSDouble_Official.txt came from `KSDoubleOut 1024 64 1e-7 20 6 1`
    (testProcedure.py:51-57).  The same run at dt=1e-3 must print the same
    header and times and converge to it (RMSE ~3e-8 at t=6, O(dt^2))."""
    r = _run(sw, 1, 0, [1024, 64, 0.001, 20, 6, 1, "ks.dat", "ks.txt"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "GPU only Swept scheme" in r.stdout
    got = open(tmp_path / "ks.dat").read().split("\n")
    ref = open(os.path.join(GOLDEN, "KSDouble_Official.txt")).read().split("\n")
    assert got[0] == ref[0] == "512 1024 0.5 "
    assert got[1] == ref[1]  # t = 0 row, identical text
    gt = [l.split()[1] for l in got[1:] if l.strip()]
    rt = [l.split()[1] for l in ref[1:] if l.strip()]
    assert gt == rt == ["0", "6", "12", "18", "20"]
    bounds = [0, 4.0e-8, 1.7e-8, 5e-9, 3.3e-9]
    for i in range(1, 5):
        g = np.array(got[1 + i].split()[2:], float)
        f = np.array(ref[1 + i].split()[2:], float)
        assert np.sqrt(np.mean((g - f) ** 2)) < bounds[i]
    line = open(tmp_path / "ks.txt").read().strip().split("\t")
    assert line[:2] == ["1024", "64"] and float(line[2]) > 0


def _oracle_states(orc, eq, prec, n, dt, steps_list):
    dtype = np.float64 if prec == 0 else np.float32
    if eq == 0:
        u = orc.heat_ic_ref(n, dtype)
        fo, fo2 = orc.heat_consts(dt, dtype)
        run = lambda s, k: orc.heat_run(s, k, fo, fo2, 0)
    elif eq == 1:
        u = orc.ks_ic_ref(n, dtype)
        c = orc.ks_consts(dt, dtype)
        run = lambda s, k: orc.ks_run(s, k, c)
    else:
        u = orc.euler_ic_ref(n, dtype)
        kk, bd = orc.euler_consts(dt, n, dtype)
        run = lambda s, k: orc.euler_run(s, k, kk, bd)
    out, done = [], 0
    for s in steps_list:
        u = run(u, s - done)
        done = s
        out.append(u.copy())
    return out


CLI_CASES = [(eq, prec, alg, 128) for eq in (0, 1, 2) for prec in (0, 1) for alg in (0, 1, 2)]
# tpb 64, the smallest the reference's timing sweep (runTest.sh) runs: 64-point
# nodes, one point per lane, for the swept and hybrid/register algorithms
CLI_CASES += [(eq, prec, alg, 64) for eq in (0, 1, 2) for prec in (0, 1) for alg in (1, 2)]


@pytest.mark.parametrize("eq,prec,alg,tpb", CLI_CASES)
def test_program_rows_equal_oracle(sw, orc, tmp_path, eq, prec, alg, tpb):
    n = 1024
    if eq == 0:
        dt, tf, freq = 0.001, 1.5, 0.5
    elif eq == 1:
        dt, tf, freq = 0.005, 4.0, 1.5
    else:
        dt, tf, freq = 1e-4, 0.03, 0.012
    r = _run(sw, eq, prec, [n, tpb, dt, tf, freq, alg, "o.dat", "t.txt"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    header, rows = _rows(tmp_path / "o.dat")
    # schedule the library replays (Heat:483-682, KS:609-851, Euler:855-1098)
    kinds = {0: (0, 1, 1), 1: (2, 3, 4), 2: (5, 6, 6)}
    kind = kinds[eq][alg]
    dts = float(np.float32(dt)) if (eq == 1 and prec == 1) else dt
    steps, times = sw.schedule(kind, dt, tf, freq, tpb, dts)
    states = _oracle_states(orc, eq, prec, n, dt, steps)
    nvar = 4 if eq == 2 else 1
    body = rows[nvar:]  # skip the t = 0 row(s)
    assert len(body) == nvar * len(steps)
    for i, (t, st) in enumerate(zip(times, states)):
        rr = body[nvar * i: nvar * (i + 1)]
        if eq == 2:
            rho, vel, en, pr = orc.euler_outputs(st, orc.euler_consts(dt, n, st.dtype)[0][1])
            want = [rho[1:-1], vel[1:-1], en[1:-1], pr[1:-1]]
            names = ["Density", "Velocity", "Energy", "Pressure"]
        else:
            want = [st]
            names = ["Temperature" if eq == 0 else "Velocity"]
        for row, name, w in zip(rr, names, want):
            tok = row.split()
            assert tok[0] == name and tok[1] == _fmt(t)
            assert tok[2:] == [_fmt(v) for v in w], (name, t)
    # row text conventions of the reference
    if eq == 0:
        lead = " Temperature " if alg == 0 else "Temperature "
        assert rows[1].startswith(lead) if len(rows) > 2 else True
        assert not open(tmp_path / "o.dat").read().endswith("\n")  # Heat:821-824
    if eq == 1:
        assert rows[-1].startswith(" Velocity ")
    assert open(tmp_path / "t.txt").read().startswith(f"{n}\t{tpb}\t")


def test_header_formats(sw, tmp_path):
    r = _run(sw, 0, 1, [1024, 64, 0.001, 0.1, 1, 0, "h.dat"], tmp_path)
    assert r.returncode == 0
    # single precision: REAL ds=0.001f printed with 10 digits (Heat:58, 763)
    assert open(tmp_path / "h.dat").read().split("\n")[0] == "%.10g %d %.10g " % (
        float(np.float32(0.001)) * 1023, 1024, float(np.float32(0.001)))
    r = _run(sw, 2, 0, [1024, 64, 1e-4, 0.001, 1, 1, "e.dat"], tmp_path)
    assert r.returncode == 0
    assert open(tmp_path / "e.dat").read().split("\n")[0] == "1 1022 %.10g " % (1.0 / 1022)
    assert "Euler --- #Blocks: 16 | Length: 1 | Precision: Double" in r.stdout


def test_error_exits(sw, tmp_path):
    r = _run(sw, 1, 0, [1024, 64, 0.001], tmp_path)
    assert r.returncode == 255 and "The Program takes" in r.stdout
    r = _run(sw, 0, 0, [1025, 64, 0.001, 1, 1, 1, "x.dat"], tmp_path)  # odd nX (Heat:735)
    assert r.returncode == 255 and "INVALID NUMERIC INPUT" in r.stdout
    r = _run(sw, 0, 0, [1024, 48, 0.001, 1, 1, 1, "x.dat"], tmp_path)  # tpb % 32
    assert r.returncode == 255 and "INVALID NUMERIC INPUT" in r.stdout
    r = _run(sw, 2, 0, [1024, 64, 0.001, 1, 1, 0, "x.dat"], tmp_path)  # dt/dx = 1.022
    assert r.returncode == 255 and "too high" in r.stdout


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("tpb", [64, 256, 1024])
def test_ks_register_self_test(sw, tmp_path, prec, tpb):
    """alg 3: the register engine (128-point nodes in VGPRs) against the
    shared engine (node width tpb) phase by phase, as KS:682-750 compares the
    register kernels with the shared ones launch by launch: both engines run
    swept tiles only (no classic step), the "WHOLE" phase puts diamonds in
    both, every phase prints EQUAL!, and the program exits -1."""
    r = _run(sw, 1, prec, [4096, tpb, 0.001, 1, 1, 3, "x.dat"], tmp_path)
    assert r.returncode == 255, r.stdout
    out = r.stdout
    assert out.count("EQUAL!") == 2 and "NOT EQUAL" not in out
    assert "---- TEST UP/DOWN" in out and "---- TEST WHOLE" in out
    counts = {}
    for line in out.splitlines():
        if line.startswith(("shared node", "register node")):
            head, rest = line.split(":")
            v = rest.split()
            counts.setdefault(head.split()[0], []).append(dict(zip(v[::2], map(int, v[1::2]))))
    assert set(counts) == {"shared", "register"}
    for engine, phases in counts.items():
        updown, whole = phases
        for ph in phases:
            assert ph["classic"] == 0 and ph["up"] == 1 and ph["down"] == 1, (engine, ph)
        assert whole["diamond"] + whole["split"] >= 2, (engine, whole)
    # the engine with the taller tile runs no diamond in UP/DOWN, only in WHOLE
    assert min(p[0]["diamond"] + p[0]["split"] for p in counts.values()) == 0