"""A KS plan far past 32-bit indexing: nX = 3 * 2^30 points (24 GiB per FP64
state array, 96 GiB of plan buffers in HBM), so that point indices above
2^31 - 1 reach every kernel of an advance: the up triangle, the
boundary-straddling diamond, a node-aligned diamond, the down triangle and
classic steps.  The state stays on the GPU (torch tensors handed over as
device pointers); windows of the result are checked bit for bit against the
oracle run on the same window plus a margin wider than the stencil's reach
over the run (4 points per step), at the periodic seam, across index 2^31,
at the straddling node and in the middle."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 3 << 30
W = 1024
STEPS = 3 * (W // 8) + 7
MARGIN = 4 * STEPS + 64
DT = 0.005
CHUNK = 1 << 27


def _fill_ic(torch, u):
    """u[i] = 2 cos(19 pi x / 128) + 1e-3 N(0, 1), x = 0.5 i, chunk by chunk
    (the values only need to be reproducible within this test: the windows
    the oracle sees are read back from u itself)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(20261203)
    a = 19.0 * math.pi / 128.0 * 0.5
    for s in range(0, u.numel(), CHUNK):
        e = min(s + CHUNK, u.numel())
        x = torch.arange(s, e, dtype=torch.float64, device="cuda")
        x.mul_(a).cos_().mul_(2.0)
        x.add_(torch.randn(e - s, dtype=torch.float64, device="cuda", generator=g), alpha=1e-3)
        u[s:e].copy_(x)
        del x
    torch.cuda.synchronize()


def _window(t, lo, hi):
    """Host copy of t[lo:hi] with indices taken modulo N."""
    idx = [(lo + i) % N for i in range(0, hi - lo)]
    # contiguous runs: at most one wrap
    cut = next((k for k in range(1, len(idx)) if idx[k] != idx[k - 1] + 1), None)
    if cut is None:
        return t[idx[0]: idx[-1] + 1].cpu().numpy()
    return np.concatenate([t[idx[0]: idx[cut - 1] + 1].cpu().numpy(), t[idx[cut]: idx[-1] + 1].cpu().numpy()])


def test_ks_beyond_32bit_indices(sw, orc):
    torch = pytest.importorskip("torch")
    free, _ = torch.cuda.mem_get_info()
    need = 4 * N * 8 + N * 8 + 4 * CHUNK * 8
    if free < need:
        pytest.skip(f"needs {need / 2**30:.0f} GiB of free device memory, {free / 2**30:.0f} GiB free")
    print(f"KS nX={N} W={W} steps={STEPS}", flush=True)
    u = torch.empty(N, dtype=torch.float64, device="cuda")
    _fill_ic(torch, u)
    centres = [0, N // 2, (1 << 31) - 1, (1 << 31) + 5, N - W // 2, N - 1]
    half = 2048
    before = {c: _window(u, c - half - MARGIN, c + half + MARGIN) for c in centres}
    with sw.Plan(sw.KS, N, DT, algorithm=sw.SWEPT, tpb=W, node_width=W) as p:
        assert p.info()["tile_steps"] == W // 8
        p.set_profiling(True)
        p.set_state_device(u.data_ptr())
        p.advance(STEPS)
        p.synchronize()
        counts = {k: p.kernel_stats(k)[0] for k in (sw.K_UP, sw.K_DIAMOND, sw.K_SPLIT, sw.K_DOWN, sw.K_CLASSIC)}
        p.get_state_device(u.data_ptr())
    torch.cuda.synchronize()
    assert counts[sw.K_UP] == 1 and counts[sw.K_DOWN] == 1
    assert counts[sw.K_DIAMOND] == 1 and counts[sw.K_SPLIT] == 1
    assert counts[sw.K_CLASSIC] == 2 * 7  # two stage launches per classic step
    coef = orc.ks_consts(DT, np.float64)
    for c in centres:
        want = orc.ks_run(before[c], STEPS, coef)[MARGIN: MARGIN + 2 * half]
        got = _window(u, c - half, c + half)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, f"window at {c}: {bad.size} mismatches, first at offset {bad[:8] - half}"
    assert torch.isfinite(u).all().item()
    del u
    torch.cuda.empty_cache()
