"""The reference's time bookkeeping (sw1d_schedule, C++) — checked against the
times printed in the reference's own fixture and against a transcription of
the wrapper loops (Heat:483-510, 541-682; KS:609-631, 676-851;
Euler:855-888, 924-1098).  CPU only: the schedule is host code."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def test_fixture_times_from_ks_swept_schedule(sw):
    """KSDouble_Official.txt: KS swept, nX=1024, tpb=64, dt=1e-7, tf=20,
    freq=6 (testProcedure.py:51-57) printed rows at 6, 12, 18, 20."""
    steps, times = sw.schedule(sw.SCHED_KS_SWEPT, 1e-7, 20.0, 6.0, 64, 1e-7)
    printed = ["%.10g" % t for t in times]
    rows = [l.split()[1] for l in open(os.path.join(GOLDEN, "KSDouble_Official.txt"))][1:]
    assert printed == rows[1:]
    assert all(s % 16 == 0 for s in steps)  # whole cycles of tpb/4 steps


def _restated(kind, dt, tf_arg, freq, tpb, dts):
    """Line-by-line transcription of the reference wrapper loops."""
    t_end = tf_arg - 0.25 * dt
    twrite = freq - 0.25 * dts
    rows = []
    if kind == 0:  # heat classic
        full = dts + dts
        n, t = 2, full
        while t < t_end:
            n += 2
            t += full
            if t > twrite:
                rows.append((n, t))
                t += full
                twrite += freq
    elif kind in (2, 5):  # KS / Euler classic
        n, t = 0, 0.0
        cond = (lambda a, b: a <= b) if kind == 2 else (lambda a, b: a < b)
        while cond(t, t_end):
            n += 1
            t += dts
            if t > twrite:
                rows.append((n, t))
                twrite += freq
    else:
        if kind == 1:
            full, per = dts * tpb, tpb
        elif kind == 4:
            full, per = 0.25 * dts * 32.0, 8
        else:
            full, per = 0.25 * dts * tpb, tpb // 4
        n, t = per, full
        while t < t_end:
            n += per
            t += full
            if t > twrite:
                rows.append((n, t))
                n += per
                t += full
                twrite += freq
    rows.append((n, t))
    return rows


@pytest.mark.parametrize("kind", range(7))
@pytest.mark.parametrize("dt,tf,freq,tpb", [
    (1e-3, 1.0, 0.3, 64), (5e-3, 20.0, 6.0, 128), (1e-4, 0.22, 0.066, 256),
    (2.5e-4, 3.0, 7.0, 32), (1e-5, 0.05, 0.01, 1024),
])
def test_schedule_matches_wrapper_loops(sw, kind, dt, tf, freq, tpb):
    dts = float(np.float32(dt)) if kind in (2, 3, 4) else dt
    steps, times = sw.schedule(kind, dt, tf, freq, tpb, dts)
    want = _restated(kind, dt, tf, freq, tpb, dts)
    assert list(zip(steps, times)) == want


def test_heat_classic_quirk_clock_jump(sw):
    """Heat:506 advances the clock after every output without computing."""
    steps, times = sw.schedule(sw.SCHED_HEAT_CLASSIC, 0.01, 1.0, 0.25, 64)
    assert len(steps) > 2
    # printed time runs ahead of steps*dt by 2 dt per earlier output
    for i, (s, t) in enumerate(zip(steps[:-1], times[:-1])):
        assert abs(t - (s + 2 * i) * 0.01) < 1e-9


def test_schedule_capacity_error(sw):
    import ctypes as C

    cnt = C.c_int64(0)
    st = (C.c_int64 * 1)()
    tm = (C.c_double * 1)()
    rc = sw.lib().sw1d_schedule(3, 1e-3, 1.0, 0.1, 64, 1e-3, st, tm, 1, C.byref(cnt))
    assert rc == -1 and cnt.value > 1
