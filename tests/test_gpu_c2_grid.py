"""The KS timing sweep's grid (the reference's runTest.sh / testProcedure.py
configuration: dx = 0.5, dt = 0.005, the noisy cosine initial condition,
nX in {2^14, 2^16, 2^18, 2^20}, tpb 64..1024, FP64 and FP32, algorithms
0 (classic), 1 (shared swept) and 2 (register engine)) at its full 20000
steps, every case against the OpenMP oracle bit for bit.  At 2^14 and 2^16
every node width runs; at 2^18 and 2^20 the widths the bench and the tuned
table use.  20000 steps is 156 diamond rows at node width 1024 and 2500 at
64, plus a remainder of classic steps wherever the tile height does not
divide 20000."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 20000
DT = 0.005
THREADS = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))

CASES = []
for lg in (14, 16):
    for prec in (0, 1):
        CASES.append((lg, prec, 0, 256))
        CASES += [(lg, prec, 1, W) for W in (64, 128, 256, 512, 1024)]
        CASES.append((lg, prec, 2, 128))
for lg in (18, 20):
    CASES += [(lg, 0, 0, 256), (lg, 0, 1, 512), (lg, 0, 1, 1024), (lg, 1, 1, 1024), (lg, 0, 2, 128)]

_WANT = {}


def _want(orc, lg, prec):
    """The oracle's state after STEPS steps, computed once per size and precision."""
    key = (lg, prec)
    if key not in _WANT:
        dtype = np.float64 if prec == 0 else np.float32
        u0 = orc.ic_ks_noisy(1 << lg, 7).astype(dtype)
        _WANT.clear()  # one entry at a time (2^20 FP64 is 8 MB, the cases come grouped)
        _WANT[key] = (u0, orc.ks_run(u0, STEPS, orc.ks_consts(DT, dtype), THREADS))
    return _WANT[key]


@pytest.mark.parametrize("lg,prec,alg,W", CASES)
def test_c2_grid_matches_oracle(sw, orc, lg, prec, alg, W):
    n = 1 << lg
    print(f"C2 nX=2^{lg} prec={prec} alg={alg} W={W} steps={STEPS}", flush=True)
    sys.stdout.flush()
    u0, want = _want(orc, lg, prec)
    kw = dict(precision=prec, algorithm=alg, tpb=W)
    if alg != 0:
        kw["node_width"] = W
    with sw.Plan(sw.KS, n, DT, **kw) as p:
        if alg != 0:
            assert p.info()["node_width"] == W
        p.set_state(u0)
        p.advance(STEPS)
        got = p.get_state()
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:8]}"
    assert np.all(np.isfinite(got)) and not np.array_equal(got, u0)
