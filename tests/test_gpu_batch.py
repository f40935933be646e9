"""Batched persistent solve: every problem of a mixed heat/KS batch of uneven
sizes is bitwise the oracle's classic result (periodic BCs)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    import importlib

    return importlib.import_module("1dsweptcuda_amd.workloads")


@pytest.mark.parametrize("waves", [2, 4, 8])
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nsteps", [203, 40])
def test_batch_matches_oracle(sw, orc, wl, prec, nsteps, waves):
    count = 40
    eq, nx, dt, dx = wl.batch_problems(count, seed=prec + 5, lo=7.0, hi=13.0)
    # every tile width, also wider than the problem (windows wrap several times)
    tps = np.array([(128, 256, 512, 1024)[i % 4] for i in range(count)], np.int32)
    dtype = np.float64 if prec == 0 else np.float32
    ics = [u.astype(dtype) for u in wl.batch_ics(eq, nx, dx)]
    print(f"batch prec={prec} steps={nsteps} waves={waves} sizes={nx.min()}..{nx.max()}", flush=True)
    with sw.Batch(eq, nx, dt, dx, precision=prec, tile_points=tps, waves=waves) as b:
        b.set_state(np.concatenate(ics))
        b.advance(nsteps)
        got = b.split(b.get_state())
    for i in range(count):
        if eq[i] == 0:
            fo, fo2 = orc.heat_consts(dt[i], dtype, ds=dx[i])
            want = orc.heat_run(ics[i], nsteps, fo, fo2, 1)
        else:
            want = orc.ks_run(ics[i], nsteps, orc.ks_consts(dt[i], dtype, dx=dx[i]))
        assert np.array_equal(got[i], want), (i, int(eq[i]), int(nx[i]), int(tps[i]))


def test_batch_replay_and_repeat(sw, orc, wl):
    """Two advances of k steps (graph replayed) equal one oracle run of 2k."""
    eq, nx, dt, dx = wl.batch_problems(12, seed=3, lo=8.0, hi=11.0)
    ics = wl.batch_ics(eq, nx, dx)
    with sw.Batch(eq, nx, dt, dx) as b:
        b.set_state(np.concatenate(ics))
        b.advance(64)
        b.advance(64)
        got = b.split(b.get_state())
    for i in range(12):
        if eq[i] == 0:
            fo, fo2 = orc.heat_consts(dt[i], ds=dx[i])
            want = orc.heat_run(ics[i], 128, fo, fo2, 1)
        else:
            want = orc.ks_run(ics[i], 128, orc.ks_consts(dt[i], dx=dx[i]))
        assert np.array_equal(got[i], want), i


def test_batch_workgroup_setting(sw, orc, wl):
    """Waves per workgroup: only 2, 4, 8; changing it between advances
    rebuilds the launch and keeps the bits."""
    eq, nx, dt, dx = wl.batch_problems(10, seed=9, lo=8.0, hi=12.0)
    ics = wl.batch_ics(eq, nx, dx)
    print("batch workgroup switch", flush=True)
    with sw.Batch(eq, nx, dt, dx) as b:
        with pytest.raises(sw.SweptError):
            b.set_workgroup(3)
        b.set_state(np.concatenate(ics))
        b.advance(48)
        b.set_workgroup(8)
        b.advance(48)
        b.set_workgroup(2)
        b.advance(48)
        got = b.split(b.get_state())
    for i in range(10):
        if eq[i] == 0:
            fo, fo2 = orc.heat_consts(dt[i], ds=dx[i])
            want = orc.heat_run(ics[i], 144, fo, fo2, 1)
        else:
            want = orc.ks_run(ics[i], 144, orc.ks_consts(dt[i], dx=dx[i]))
        assert np.array_equal(got[i], want), i


def _stratified(eq, nx, per_class=6):
    """Indices covering every size class floor(log2 nX) and both equations,
    `per_class` each (the largest of the class first), plus the largest and
    smallest problem of the whole batch."""
    cls = np.floor(np.log2(nx)).astype(int)
    pick = {int(np.argmax(nx)), int(np.argmin(nx))}
    for c in np.unique(cls):
        for e in (0, 1):
            idx = np.flatnonzero((cls == c) & (eq == e))
            idx = idx[np.argsort(-nx[idx], kind="stable")]
            pick.update(int(i) for i in idx[:per_class])
    return sorted(pick)


@pytest.fixture(scope="module")
def c4_workload(wl):
    eq, nx, dt, dx = wl.batch_problems(4096, seed=2024)
    return eq, nx, dt, dx, wl.batch_ics(eq, nx, dx)


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_workload_full_shape(sw, orc, wl, c4_workload, prec):
    """The batched workload at its own shape: 4096 heat/KS problems of
    nX = floor(2^U[10,16]), 5000 steps, the tuned tile widths and waves per
    workgroup, in one persistent launch.  The grid holds fewer workgroups than
    problems, so workgroups pull further problems from the queue.  A sample
    covering every size class and both equations (>= 64 problems, with the
    largest) is bitwise the oracle's."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    import importlib

    at = importlib.import_module("1dsweptcuda_amd.autotune")
    eq, nx, dt, dx, ics64 = c4_workload
    steps = 5000
    dtype = np.float64 if prec == 0 else np.float32
    table = at.load()
    assert table, "the tuned tile table is missing"
    tps = at.tile_points_for(nx, table)
    waves = at.load_waves()
    ics = [u.astype(dtype) for u in ics64]
    print(f"batch C4 prec={prec} problems={nx.size} points={int(nx.sum())} steps={steps} waves={waves}",
          flush=True)
    with sw.Batch(eq, nx, dt, dx, precision=prec, tile_points=tps, waves=waves) as b:
        info = b.info()
        assert info["problems"] == 4096 and info["total_points"] == int(nx.sum())
        assert info["workgroups"] < nx.size
        b.set_state(np.concatenate(ics))
        b.advance(steps)
        got = b.split(b.get_state())
    sample = _stratified(eq, nx)
    assert len(sample) >= 64 and int(np.argmax(nx)) in sample
    assert set(np.floor(np.log2(nx[sample])).astype(int)) == set(range(10, 16))

    def oracle(i):
        if eq[i] == 0:
            fo, fo2 = orc.heat_consts(dt[i], dtype, ds=dx[i])
            return orc.heat_run(ics[i], steps, fo, fo2, 1)
        return orc.ks_run(ics[i], steps, orc.ks_consts(dt[i], dtype, dx=dx[i]))

    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(workers) as pool:
        wants = list(pool.map(oracle, sample))
    bad = [(i, int(eq[i]), int(nx[i])) for i, w in zip(sample, wants) if not np.array_equal(got[i], w)]
    assert not bad, bad


def test_direct_launches_without_graphs(sw, tmp_path):
    """SW1D_LAUNCH_GRAPHS=0 (read once per process, so run in a child process):
    trapezoid plans, classic plans and batches launch their kernels one by one
    instead of replaying hipGraphs, with the same bits."""
    import os
    import subprocess
    import sys

    from conftest import ORACLE_DIR, ROOT

    script = r'''
import importlib, sys
import numpy as np
sys.path[:0] = [sys.argv[1], sys.argv[2]]
import pyoracle as orc
sw = importlib.import_module("1dsweptcuda_amd")
wl = importlib.import_module("1dsweptcuda_amd.workloads")
n, dt = 4096, 0.005
u0 = orc.ic_ks_noisy(n, 3)
want = orc.ks_run(u0, 700, orc.ks_consts(dt))
for alg, kw in ((sw.TRAPEZOID, dict(node_width=512)), (sw.CLASSIC, {})):
    with sw.Plan(sw.KS, n, dt, algorithm=alg, **kw) as p:
        p.set_state(u0)
        p.advance(300)
        p.advance(400)
        assert np.array_equal(p.get_state(), want), alg
eq, nx, dts, dx = wl.batch_problems(24, seed=4, lo=8.0, hi=12.0)
ics = wl.batch_ics(eq, nx, dx)
with sw.Batch(eq, nx, dts, dx) as b:
    b.set_state(np.concatenate(ics))
    b.advance(64)
    b.advance(100)
    got = b.split(b.get_state())
for i in range(24):
    if eq[i] == 0:
        fo, fo2 = orc.heat_consts(dts[i], ds=dx[i])
        w = orc.heat_run(ics[i], 164, fo, fo2, 1)
    else:
        w = orc.ks_run(ics[i], 164, orc.ks_consts(dts[i], dx=dx[i]))
    assert np.array_equal(got[i], w), i
print("direct launches ok")
'''
    print("direct launches (SW1D_LAUNCH_GRAPHS=0) in a child process", flush=True)
    env = dict(os.environ, SW1D_LAUNCH_GRAPHS="0")
    r = subprocess.run([sys.executable, "-c", script, ROOT, ORACLE_DIR], env=env, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "direct launches ok" in r.stdout
