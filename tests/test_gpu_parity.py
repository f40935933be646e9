"""GPU parity: every algorithm of libswept1d against the CPU oracle, bitwise.

The reference's invariant is swept == classic (ResultPlots/ExactTesting/
consistency.out); the build holds every path — classic, swept, hybrid /
register — to the same bits as the serial oracle, which at the chaotic KS
horizon is the only way to stay within 1e-12 relative.  Calls go through the
C ABI (ctypes on lib/libswept1d.so)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ks_u0(orc, n, dtype):
    return orc.ic_ks_noisy(n, 11).astype(dtype)


def _oracle(orc, eq, u0, nsteps, dt, dtype, bc, n):
    if eq == 0:
        fo, fo2 = orc.heat_consts(dt, dtype)
        return orc.heat_run(u0, nsteps, fo, fo2, 0 if bc != 2 else 1)
    if eq == 1:
        return orc.ks_run(u0, nsteps, orc.ks_consts(dt, dtype))
    k, bd = orc.euler_consts(dt, n, dtype)
    return orc.euler_run(u0, nsteps, k, bd)


def _ic(orc, eq, n, dtype, bc):
    if eq == 0:
        return orc.heat_ic_ref(n, dtype) if bc != 2 else orc.ic_heat_fourier(n, 5)[0].astype(dtype)
    if eq == 1:
        return _ks_u0(orc, n, dtype)
    return orc.euler_ic_ref(n, dtype)


def _dt(eq, n, bc):
    if eq == 0:
        return 0.001 if bc != 2 else 0.0025
    if eq == 1:
        return 0.005
    return 0.1 / (n - 2)


def test_device_lane_shift_check(sw):
    _say("device_check")
    assert sw.device_check()


CASES = []
for eq, bc in ((0, 0), (0, 2), (1, 0), (2, 0)):
    for prec in (0, 1):
        for alg in (0, 1, 2):
            CASES.append((eq, bc, prec, alg))


@pytest.mark.parametrize("eq,bc,prec,alg", CASES)
def test_parity_vs_oracle(sw, orc, eq, bc, prec, alg):
    n = 4096
    dtype = np.float64 if prec == 0 else np.float32
    dt = _dt(eq, n, bc)
    tpb = 256
    nsteps = 300 if eq != 2 else 200  # not a multiple of the tile height: remainder path too
    _say(f"parity eq={eq} bc={bc} prec={prec} alg={alg} n={n} tpb={tpb} steps={nsteps}")
    u0 = _ic(orc, eq, n, dtype, bc)
    with sw.Plan(eq, n, dt, precision=prec, algorithm=alg, tpb=tpb, bc=bc) as p:
        p.set_state(u0)
        p.advance(nsteps // 3)
        p.advance(nsteps - nsteps // 3)
        got = p.get_state()
    want = _oracle(orc, eq, u0, nsteps, dt, dtype, bc, n)
    assert got.dtype == want.dtype
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:8]}: {got[bad[:4]]} vs {want[bad[:4]]}"


@pytest.mark.parametrize("W", [512, 1024])
def test_ks_phased_diamond_single_precision(sw, orc, W):
    """The KS diamonds with P >= 8 points per lane run in four region widths
    from an LDS image of the node; FP32 through every launch kind."""
    n = 8 * W
    dt = _dt(1, n, 1)
    nsteps = 5 * (W // 8) + 3
    _say(f"phased KS FP32 W={W} n={n} steps={nsteps}")
    u0 = _ic(orc, 1, n, np.float32, 1)
    with sw.Plan(1, n, dt, precision=1, algorithm=1, node_width=W) as p:
        p.set_state(u0)
        p.advance(nsteps)
        got = p.get_state()
    assert np.array_equal(got, _oracle(orc, 1, u0, nsteps, dt, np.float32, 0, n))


NODE_CASES = [(eq, W, prec, alg) for eq in (0, 1, 2) for W in (64, 128, 256, 512, 1024)
              for prec in (0, 1) for alg in (1, 2) if not (eq == 2 and W > 512)]  # Euler nodes stop at 512


@pytest.mark.parametrize("eq,W,prec,alg", NODE_CASES)
def test_node_widths(sw, orc, eq, W, prec, alg):
    """Every node width the equation allows, 64 included (one point per lane,
    where a radius-2 stencil's outer neighbour comes from two lanes away), in
    both precisions, with the shared swept algorithm and with algorithm 2 at
    the same width (heat and Euler: the hybrid, whose boundary node the host
    computes; KS: the register engine)."""
    n = 8 * W
    dtype = np.float64 if prec == 0 else np.float32
    dt = _dt(eq, n, 0)
    H = W // 2 if eq == 0 else W // 8
    nsteps = 5 * H + 3
    _say(f"node width eq={eq} W={W} prec={prec} alg={alg} n={n} steps={nsteps}")
    u0 = _ic(orc, eq, n, dtype, 0)
    with sw.Plan(eq, n, dt, precision=prec, algorithm=alg, tpb=W, node_width=W) as p:
        assert p.info()["node_width"] == W
        p.set_state(u0)
        p.advance(nsteps)
        got = p.get_state()
    want = _oracle(orc, eq, u0, nsteps, dt, dtype, 0, n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("eq,lg,prec", [(0, 14, 0), (1, 14, 0), (1, 18, 0), (2, 16, 0),
                                         (1, 14, 1), (1, 18, 1), (2, 14, 1), (0, 16, 1)])
def test_auto_node_width(sw, orc, eq, lg, prec):
    """node_width="auto" takes the autotuner's width for the equation, precision
    and size class (tuned/node_widths.json) and gives the oracle's bits, as
    every width does."""
    import importlib
    at = importlib.import_module("1dsweptcuda_amd.autotune")
    n = 1 << lg
    W = at.node_width_for(eq, n, prec)
    assert W > 0, "no tuned node width for this equation and precision"
    dtype = np.float64 if prec == 0 else np.float32
    dt = _dt(eq, n, 0)
    H = W // 2 if eq == 0 else W // 8
    nsteps = 2 * H + 5
    _say(f"auto node width eq={eq} n={n} prec={prec} W={W} steps={nsteps}")
    u0 = _ic(orc, eq, n, dtype, 0)
    with sw.Plan(eq, n, dt, precision=prec, algorithm=1, node_width="auto") as p:
        assert p.info()["node_width"] == W
        p.set_state(u0)
        p.advance(nsteps)
        got = p.get_state()
    want = _oracle(orc, eq, u0, nsteps, dt, dtype, 0, n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("eq,alg,W", [(0, 1, 128), (1, 1, 128), (2, 1, 128), (1, 2, 0), (0, 3, 256),
                                      (1, 3, 128), (2, 3, 128)])
def test_tile_graph_blocks(sw, orc, eq, alg, W):
    """Long advances from either buffer: trapezoid launches (alg 3) replay
    blocks of 16 launches from a hipGraph per starting buffer, swept diamonds
    (alg 1, and KS's register alg 2) go one by one.  Two advances of 37 tile
    launches and a remainder each, the second starting from the other buffer,
    give the oracle's bits, with per-kernel profiling off and on (which
    launches one by one)."""
    n = 1024
    dtype = np.float64
    dt = _dt(eq, n, 0)
    u0 = _ic(orc, eq, n, dtype, 0)
    results = []
    for profiling in (False, True):
        with sw.Plan(eq, n, dt, algorithm=alg, node_width=W, tpb=W or 64) as p:
            per = p.info()["tile_steps"]
            nsteps = 37 * per + 5
            _say(f"tile graph eq={eq} alg={alg} W={p.info()['node_width']} steps=2x{nsteps} "
                 f"profiling={profiling}")
            p.set_profiling(profiling)
            p.set_state(u0)
            p.advance(nsteps)
            p.advance(nsteps)
            results.append(p.get_state())
    want = _oracle(orc, eq, u0, 2 * nsteps, dt, dtype, 0, n)
    assert np.array_equal(results[0], want)
    assert np.array_equal(results[1], want)


def test_node_width_argument_validation(sw):
    with pytest.raises(ValueError):
        sw.Plan(1, 1 << 14, 0.005, node_width="widest")


TRAP_CASES = []
for eq, bc in ((0, 0), (0, 2), (1, 0), (2, 0)):
    for prec in (0, 1):
        for W, depth in ((128, 1), (256, 1), (512, 1), (1024, 2)):
            if eq == 2 and W > 256:
                continue
            TRAP_CASES.append((eq, bc, prec, W, depth))


@pytest.mark.parametrize("eq,bc,prec,W,depth", TRAP_CASES)
def test_trapezoid_parity(sw, orc, eq, bc, prec, W, depth):
    # periodic problems take any nx (last tile clipped); the others nx = k*W
    n = 4099 if (eq == 1 or bc == 2) else 4096
    if W >= 1024:
        n = 3 * W + (5 if (eq == 1 or bc == 2) else 0)
    dtype = np.float64 if prec == 0 else np.float32
    dt = _dt(eq, n, bc)
    nsteps = 97 if eq != 0 else 301
    _say(f"trapezoid eq={eq} bc={bc} prec={prec} W={W} depth={depth} n={n} steps={nsteps}")
    u0 = _ic(orc, eq, n, dtype, bc)
    with sw.Plan(eq, n, dt, precision=prec, algorithm=3, node_width=W, tile_depth=depth, bc=bc) as p:
        p.set_state(u0)
        p.advance(nsteps)
        got = p.get_state()
    want = _oracle(orc, eq, u0, nsteps, dt, dtype, bc, n)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:8]}"


@pytest.mark.parametrize("eq", [0, 1, 2])
def test_classic_graph_blocks(sw, orc, eq):
    """The classic algorithm replays blocks of 64 steps from a hipGraph (one
    per starting buffer for heat, which alternates buffers); advancing in
    pieces that start on either buffer gives the oracle's bits."""
    n = 4096 if eq != 2 else 2048
    dtype = np.float64
    dt = _dt(eq, n, 0)
    u0 = _ic(orc, eq, n, dtype, 0)
    pieces = [3, 130, 64, 1, 200]
    _say(f"classic graph blocks eq={eq} n={n} pieces={pieces}")
    with sw.Plan(eq, n, dt, algorithm=0) as p:
        p.set_state(u0)
        for k in pieces:
            p.advance(k)
        got = p.get_state()
    want = _oracle(orc, eq, u0, sum(pieces), dt, dtype, 0, n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("alg", [0, 1, 3])
def test_periodic_heat_configuration_c1(sw, orc, alg):
    """The periodic heat case: L = 1, nX = 1024, diffusivity 8.418e-5,
    Fo = 0.25, seeded IC 1 + sum_k a_k sin(2 pi k x), 2000 steps, node width
    64 (tpb; trapezoid tiles of 128).  Classic, swept and trapezoid give the
    oracle's bits, and the oracle follows the discrete decaying-mode solution,
    each mode k damped by (1 - 4 Fo sin^2(pi k / nX)) per step."""
    n, L, alpha = 1024, 1.0, 8.418e-5
    dx = L / n
    dt = 0.25 * dx * dx / alpha
    u0, a = orc.ic_heat_fourier(n, 7, L=L)
    fo, fo2 = orc.heat_consts(dt, ds=dx)
    want = orc.heat_run(u0, 2000, fo, fo2, 1)
    x = np.arange(n) * dx
    exact = 1.0 + sum(a[k - 1] * (1.0 - 4.0 * fo * np.sin(np.pi * k / n) ** 2) ** 2000 * np.sin(2 * np.pi * k * x / L)
                      for k in range(1, 5))
    assert np.max(np.abs(want - exact)) < 1e-11
    _say(f"periodic heat C1 alg={alg}")
    with sw.Plan(0, n, dt, algorithm=alg, tpb=64, node_width=128 if alg == 3 else 0, bc=2, dx=dx) as p:
        p.set_state(u0)
        p.advance(2000)
        got = p.get_state()
    assert np.array_equal(got, want)


# (equation, bc, precision, algorithm, node width, log2 nX): at least 2048
# nodes, two per SIMD of MI355X's 1024, where the swept stages run their final
# fma under an EXEC mask instead of a compare and select (KS FP32 at W = 512:
# the diamond's stages with compile-time lane sets, one branch per set; heat
# at its tuned W = 512 with both kinds of end and in the hybrid)
XM_CASES = [(0, 0, 0, 1, 128, 18), (0, 2, 1, 1, 64, 17), (0, 0, 0, 2, 128, 18), (1, 0, 0, 1, 64, 17),
            (1, 0, 1, 1, 256, 19), (1, 0, 0, 1, 512, 20), (1, 0, 1, 1, 512, 20), (1, 0, 0, 2, 128, 18),
            (1, 0, 1, 2, 128, 18), (0, 0, 0, 1, 512, 20), (0, 0, 1, 1, 512, 20), (0, 2, 0, 1, 512, 20),
            (0, 0, 0, 2, 512, 20)]


@pytest.mark.parametrize("eq,bc,prec,alg,W,lg", XM_CASES)
def test_exec_masked_stage_form(sw, orc, eq, bc, prec, alg, W, lg):
    """Plans with two or more nodes per SIMD take the EXEC-masked form of the
    swept stages; it gives the oracle's bits like the compare-and-select form."""
    n = 1 << lg
    dtype = np.float64 if prec == 0 else np.float32
    dt = _dt(eq, n, bc)
    H = W // 2 if eq == 0 else W // 8
    nsteps = 3 * H + 5
    _say(f"exec-masked stages eq={eq} bc={bc} prec={prec} alg={alg} W={W} n={n} steps={nsteps}")
    u0 = _ic(orc, eq, n, dtype, bc)
    with sw.Plan(eq, n, dt, precision=prec, algorithm=alg, tpb=W, node_width=W, bc=bc) as p:
        assert p.info()["node_width"] == W and n // W >= 2048
        p.set_state(u0)
        p.advance(nsteps)
        got = p.get_state()
    if eq == 0:
        fo, fo2 = orc.heat_consts(dt, dtype)
        want = orc.heat_run(u0, nsteps, fo, fo2, 0 if bc != 2 else 1, 8)
    else:
        want = orc.ks_run(u0, nsteps, orc.ks_consts(dt, dtype), 8)
    assert np.array_equal(got, want)
