"""A KS plan far past 32-bit indexing: nX = 3 * 2^30 points (24 GiB per FP64
state array, 96 GiB of plan buffers in HBM), so that point indices above
2^31 - 1 reach every kernel of an advance: the up triangle, the
boundary-straddling diamond, a node-aligned diamond, the down triangle and
classic steps.  The state stays on the GPU (torch tensors handed over as
device pointers); windows of the result are checked bit for bit against the
oracle run on the same window plus a margin wider than the stencil's reach
over the run (4 points per step), at the periodic seam, across index 2^31,
at the straddling node and in the middle.

The GPU part runs in its own process that initialises torch's HIP runtime
before the library's: torch ships its own copy of the HIP runtime, and in a
process where the library's copy started first torch finds no device."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = 3 << 30
W = 1024
STEPS = 3 * (W // 8) + 7
MARGIN = 4 * STEPS + 64
HALF = 2048
CENTRES = [0, N // 2, (1 << 31) - 1, (1 << 31) + 5, N - W // 2, N - 1]

_CHILD = r"""
import importlib, json, math, sys
import numpy as np
import torch
root, out, N, W, STEPS, MARGIN, HALF = sys.argv[1], sys.argv[2], *map(int, sys.argv[3:8])
centres = json.loads(sys.argv[8])
free, _ = torch.cuda.mem_get_info()
CHUNK = 1 << 27
need = 4 * N * 8 + N * 8 + 4 * CHUNK * 8
if free < need:
    print(json.dumps({"skip": f"needs {need / 2**30:.0f} GiB of free device memory, {free / 2**30:.0f} GiB free"}))
    sys.exit(0)
sys.path.insert(0, root)
sw = importlib.import_module("1dsweptcuda_amd")
u = torch.empty(N, dtype=torch.float64, device="cuda")
g = torch.Generator(device="cuda")
g.manual_seed(20261203)
a = 19.0 * math.pi / 128.0 * 0.5  # 2 cos(19 pi x / 128) + 1e-3 N(0, 1), x = 0.5 i
for s in range(0, N, CHUNK):
    e = min(s + CHUNK, N)
    x = torch.arange(s, e, dtype=torch.float64, device="cuda")
    x.mul_(a).cos_().mul_(2.0)
    x.add_(torch.randn(e - s, dtype=torch.float64, device="cuda", generator=g), alpha=1e-3)
    u[s:e].copy_(x)
    del x
torch.cuda.synchronize()

def window(lo, hi):
    idx = np.arange(lo, hi) % N
    cut = np.flatnonzero(np.diff(idx) != 1)
    if cut.size == 0:
        return u[int(idx[0]): int(idx[-1]) + 1].cpu().numpy()
    c = int(cut[0]) + 1
    return np.concatenate([u[int(idx[0]): int(idx[c - 1]) + 1].cpu().numpy(),
                           u[int(idx[c]): int(idx[-1]) + 1].cpu().numpy()])

before = [window(c - HALF - MARGIN, c + HALF + MARGIN) for c in centres]
with sw.Plan(sw.KS, N, 0.005, algorithm=sw.SWEPT, tpb=W, node_width=W) as p:
    H = p.info()["tile_steps"]
    p.set_profiling(True)
    p.set_state_device(u.data_ptr())
    p.advance(STEPS)
    p.synchronize()
    counts = {k: p.kernel_stats(getattr(sw, k))[0] for k in ("K_UP", "K_DIAMOND", "K_SPLIT", "K_DOWN", "K_CLASSIC")}
    p.get_state_device(u.data_ptr())
torch.cuda.synchronize()
after = [window(c - HALF, c + HALF) for c in centres]
finite = bool(torch.isfinite(u).all().item())
np.savez(out, *before, *after)
print(json.dumps({"counts": counts, "finite": finite, "tile_steps": H}))
"""


def test_ks_beyond_32bit_indices(orc, tmp_path):
    import json

    print(f"KS nX={N} W={W} steps={STEPS}", flush=True)
    out = tmp_path / "windows.npz"
    r = subprocess.run([sys.executable, "-c", _CHILD, ROOT, str(out), str(N), str(W), str(STEPS), str(MARGIN),
                        str(HALF), json.dumps(CENTRES)], capture_output=True, text=True, timeout=280)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-3000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    if "skip" in res:
        pytest.skip(res["skip"])
    assert res["tile_steps"] == W // 8 and res["finite"]
    c = res["counts"]
    assert c["K_UP"] == 1 and c["K_DOWN"] == 1 and c["K_DIAMOND"] == 1 and c["K_SPLIT"] == 1
    assert c["K_CLASSIC"] == 2 * 7  # two stage launches per classic step
    z = np.load(out)
    k = len(CENTRES)
    coef = orc.ks_consts(0.005, np.float64)
    for i, centre in enumerate(CENTRES):
        before, got = z[f"arr_{i}"], z[f"arr_{k + i}"]
        want = orc.ks_run(before, STEPS, coef)[MARGIN: MARGIN + 2 * HALF]
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, f"window at {centre}: {bad.size} mismatches, first at offset {bad[:8] - HALF}"
