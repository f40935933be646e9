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
@pytest.mark.parametrize("W,nsteps", [(64, 8 * 5 + 3), (128, 16 * 5), (256, 32 * 3 + 5), (512, 64)])
def test_ks_swept_equals_classic(orc, dtype, W, nsteps):
    n = 2048
    e = Emu("ks", n, 0.005, dtype, W=W)
    u0 = _ks_ic(orc, n, dtype)
    sw = e.run(u0, nsteps)
    cl = orc.ks_run(u0, nsteps, e.coef)
    assert np.array_equal(sw, cl)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("W,nsteps", [(64, 8 * 5 + 1), (128, 16 * 7), (256, 32 * 2 + 3)])
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


@pytest.mark.parametrize("P", [4, 8, 16])
def test_ks_region_phases_contain_active_sets(P):
    """The KS diamond kernel (k_ks_swept, P >= 4) runs its stage levels on
    centred regions of 1, 2, .. P-1 points per lane below the full width
    (ks_region); the same thresholds as Euler's two-cell stencil."""
    test_euler_region_phases_contain_active_sets(P, tuple(range(1, P)))


@pytest.mark.parametrize("P", [2, 4, 8, 16])
def test_heat_region_phases_contain_active_sets(P):
    """Same check for the heat diamond kernel (k_heat_swept, P >= 2): one-point
    stencil, thresholds 2s-1, 2s+1 down and W-1-2m, W-3-2m up; regions of
    1, 2, .. P-1 points per lane."""
    W, H = 64 * P, 32 * P
    step = 1
    pcs = list(range(step, P, step))
    down = lambda s0, s1: [t for s in range(s0, s1 + 1, 2) for t in (2 * s - 1, 2 * s + 1)]
    up = lambda m0, m1: [t for m in range(m0, m1 + 1, 2) for t in (W - 1 - 2 * m, W - 3 - 2 * m)]
    phases = [(pc, down(32 * (pc - step) + 1, 32 * pc - 1)) for pc in pcs]
    phases.append((P, down(32 * (P - step) + 1, H - 1) + up(1, 32 * step - 1)))
    for pc in reversed(pcs):
        if pc == step:
            phases.append((pc, up(32 * (P - pc) + 1, H - 3) + [W - 1 - 2 * (H - 1)]))
        else:
            phases.append((pc, up(32 * (P - pc) + 1, 32 * (P - pc + step) - 1)))
    reference = down(1, H - 1) + up(1, H - 3) + [W - 1 - 2 * (H - 1)]
    assert [t for _, ts in phases for t in ts] == reference
    x = np.arange(W)
    d = np.abs(2 * x - (W - 1))
    for pc, ts in phases:
        lo, hi = W // 2 - 32 * pc, W // 2 + 32 * pc
        if pc < P:
            assert lo >= 1 and hi + 1 <= W
        halo = np.array([lo - 1, hi]) if pc < P else np.array([], int)
        for t in ts:
            act = x[d <= t]
            assert act.min() >= lo and act.max() < hi, (pc, t)
            assert np.all(d[halo] > t), (pc, t)


@pytest.mark.parametrize("P,pcs", [(4, (1, 2, 3)), (8, (1, 2, 3, 4, 5, 6, 7)), (8, (2, 3, 4, 5, 6, 7))])
def test_euler_region_phases_contain_active_sets(P, pcs):
    """k_euler_swept: two-cell stencil like KS; regions of PC = 1, 2, .. P-1
    cells per lane below the full width (kEuRegionFirst = 1), and the
    schedule from 2 that preceded it."""
    W, H = 64 * P, 16 * P
    S = [16 * pc - 1 for pc in pcs]
    M = [16 * (P - pc) + 1 for pc in pcs]
    down = lambda s0, s1: [t for s in range(s0, s1 + 1, 2) for t in (4 * s - 1, 4 * s + 3)]
    up = lambda m0, m1: [t for m in range(m0, m1 + 1, 2) for t in (W - 1 - 4 * m, W - 5 - 4 * m)]
    phases, s0 = [], 1
    for pc, s1 in zip(pcs, S):
        phases.append((pc, down(s0, s1)))
        s0 = s1 + 2
    phases.append((P, down(s0, H - 1) + up(1, M[-1] - 2)))
    for i in range(len(pcs) - 1, 0, -1):
        phases.append((pcs[i], up(M[i], M[i - 1] - 2)))
    phases.append((pcs[0], up(M[0], H - 3) + [W - 1 - 4 * (H - 1)]))
    reference = down(1, H - 1) + up(1, H - 3) + [W - 1 - 4 * (H - 1)]
    assert [t for _, ts in phases for t in ts] == reference
    x = np.arange(W)
    d = np.abs(2 * x - (W - 1))
    for pc, ts in phases:
        lo, hi = W // 2 - 32 * pc, W // 2 + 32 * pc
        if pc < P:
            assert lo >= 2 and hi + 2 <= W
        halo = np.array([lo - 2, lo - 1, hi, hi + 1]) if pc < P else np.array([], int)
        for t in ts:
            act = x[d <= t]
            assert act.min() >= lo and act.max() < hi, (pc, t)
            assert np.all(d[halo] > t), (pc, t)


def _ring_x(W, pc, lane):
    """Node-local point of `lane` in ring pc (ring_x in sw1d_kernels.hip)."""
    return W // 2 - 32 * pc + lane if lane < 32 else W // 2 + 32 * (pc - 1) + (lane - 32)


@pytest.mark.parametrize("P,first,reach", [(4, 1, 2), (8, 1, 2), (16, 1, 2),
                                           (2, 1, 1), (4, 1, 1), (16, 1, 1)])
def test_rings_partition_the_regions(P, first, reach):
    """The swept tiles move a node between HBM and its LDS image ring by ring
    (NodeRings): the loads put ring pc + 1 into the image just before region pc
    runs, and region pc's up-phase end stores ring pc.  That is sound when
    rings 1..pc are exactly region pc's points (so region pc reads only rings
    already in the image and writes no point of an outer ring), region pc's
    stencil halo lies in ring pc + 1, and rings 1..P cover the node once.
    KS (two-point reach) and heat (one-point reach) regions start at 1 point
    per lane."""
    W = 64 * P
    seen = []
    for pc in range(1, P + 1):
        ring = [_ring_x(W, pc, lane) for lane in range(64)]
        assert len(set(ring)) == 64
        seen += ring
        region = set(range(W // 2 - 32 * pc, W // 2 + 32 * pc))
        assert set(seen) == region, pc
        if first <= pc < P:
            lo, hi = W // 2 - 32 * pc, W // 2 + 32 * pc
            halo = set(range(lo - reach, lo)) | set(range(hi, hi + reach))
            nxt = {_ring_x(W, pc + 1, lane) for lane in range(64)}
            assert halo <= nxt, pc
    assert sorted(seen) == list(range(W))


@pytest.mark.parametrize("n,W", [(1024, 128), (4096, 1024), (1 << 20, 1024), (2048, 2048)])
@pytest.mark.parametrize("shifted", [False, True])
def test_ring_addresses_are_contiguous_halves(n, W, shifted):
    """A lane's ring points sit at hbase + x in the state array (NodeRings):
    node bases are multiples of W/2 and n of W, so neither half of a node
    wraps, also for the tiles centred on node boundaries (offset n - W/2),
    whose node 0 straddles the domain ends."""
    offset = n - W // 2 if shifted else 0
    for node in range(n // W):
        base = (node * W + offset) % n
        for lane in (0, 31, 32, 63):
            hbase = base if lane < 32 else (base + W // 2) % n - W // 2
            half = range(0, W // 2) if lane < 32 else range(W // 2, W)
            for x in (half[0], half[-1]):
                assert hbase + x == (base + x) % n, (node, lane, x)
