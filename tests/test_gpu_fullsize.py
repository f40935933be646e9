"""Parity at the benchmark's size, nX = 2^20: the configurations bench.py
times (KS FP64 swept at node widths 1024 and 512, trapezoid and classic, KS
FP32 swept, heat swept at 1024, Euler swept at 512) against the oracle bit
for bit, each over every launch kind (up-triangle, diamond, split diamond,
down-triangle) plus a remainder of classic steps.  The oracle runs with
OpenMP, so each case takes well under a second on the host.  A size-free
invariant is checked at the same size: a periodic KS field shifted by whole
nodes comes back shifted, bit for bit (the boundary-straddling diamond sees
different data)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1 << 20
THREADS = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(sw, eq, u0, dt, nsteps, **kw):
    with sw.Plan(eq, N, dt, **kw) as p:
        p.set_state(u0)
        p.advance(nsteps)
        return p.get_state()


@pytest.mark.parametrize("alg,W,prec", [(1, 1024, 0), (1, 512, 0), (3, 1024, 0), (0, 256, 0), (1, 512, 1)])
def test_ks_full_size_matches_oracle(sw, orc, alg, W, prec):
    dtype = np.float64 if prec == 0 else np.float32
    dt = 0.005
    nsteps = 2 * (W // 8) + 7  # up, split diamond, down, 7 classic steps (swept)
    _say(f"full-size KS alg={alg} W={W} prec={prec} steps={nsteps}")
    u0 = orc.ic_ks_noisy(N, 1).astype(dtype)
    got = _run(sw, 1, u0, dt, nsteps, precision=prec, algorithm=alg, tpb=W, node_width=W if alg != 0 else 0,
               tile_depth=2 if alg == 3 else 0)
    want = orc.ks_run(u0, nsteps, orc.ks_consts(dt, dtype), THREADS)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:8]}"


def test_heat_full_size_matches_oracle(sw, orc):
    W, dt = 1024, 0.001
    nsteps = 2 * (W // 2) + 5
    _say(f"full-size heat swept W={W} steps={nsteps}")
    u0 = orc.heat_ic_ref(N)
    got = _run(sw, 0, u0, dt, nsteps, algorithm=1, tpb=W, node_width=W)
    fo, fo2 = orc.heat_consts(dt)
    assert np.array_equal(got, orc.heat_run(u0, nsteps, fo, fo2, 0, THREADS))


def test_euler_full_size_matches_oracle(sw, orc):
    W = 512
    dt = 0.1 / (N - 2)
    nsteps = 2 * (W // 8) + 3
    _say(f"full-size Euler swept W={W} steps={nsteps}")
    q0 = orc.euler_ic_ref(N)
    got = _run(sw, 2, q0, dt, nsteps, algorithm=1, tpb=W, node_width=W)
    k, bd = orc.euler_consts(dt, N)
    assert np.array_equal(got, orc.euler_run(q0, nsteps, k, bd, THREADS))


def test_ks_full_size_shift_by_whole_nodes(sw, orc):
    W, dt, shift = 1024, 0.005, 3 * 1024
    nsteps = 3 * (W // 8)
    _say(f"full-size KS shift {shift} steps={nsteps}")
    u0 = orc.ic_ks_noisy(N, 2)
    a = _run(sw, 1, u0, dt, nsteps, algorithm=1, tpb=W, node_width=W)
    b = _run(sw, 1, np.roll(u0, shift), dt, nsteps, algorithm=1, tpb=W, node_width=W)
    assert np.array_equal(np.roll(a, shift), b)
    assert not np.array_equal(a, u0)
