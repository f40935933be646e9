"""Runs the library's hot kernels a few launches each, for hardware-counter
collection.  A counter pass (rocprofv3 --pmc) serialises and instruments every
dispatch, so this driver keeps the dispatch count small (tens per kernel, a
few hundred in all) and prints a line after each group.  Plans and batches
launch directly (SW1D_LAUNCH_GRAPHS=0): counter collection follows kernel
dispatches, not graph launches.

Groups (FP64, the bench's configurations):
  ks     KS swept, nX = 2^20, node width 1024 (up, diamonds, down)
  heat   heat swept, insulated ends, nX = 2^20, node width 512
  euler  Sod swept, nX = 2^20, node width 512
  batch  the batched workload: 4096 heat/KS problems of nX = floor(2^U[10,16]),
         5000 steps, tuned tiles and waves, FP64 then FP32 (one launch each)
  copy   a plain copy kernel of known size (torch.neg over float32), the
         reference for scaling FETCH_SIZE / WRITE_SIZE

    rocprofv3 --pmc <counters> -- python3 1dsweptcuda_amd/profile_kernels.py \\
        [--groups copy,ks,heat,euler,batch] [--tiles 10]
"""
from __future__ import annotations

import argparse
import importlib
import os
import sys
import time

os.environ["SW1D_LAUNCH_GRAPHS"] = "0"  # read once, when the first plan is created

import numpy as np  # noqa: E402

_T0 = time.perf_counter()


def _say(msg):
    print(f"[profile] {msg} at {time.perf_counter() - _T0:.1f} s", flush=True)


def _swept(sw, eq, n, W, tiles, ic, dt):
    with sw.Plan(eq, n, dt, algorithm=sw.SWEPT, tpb=W, node_width=W) as p:
        H = p.info()["tile_steps"]
        p.set_state(ic)
        p.advance(tiles * H)
        p.synchronize()
    return tiles + 1


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--groups", default="copy,ks,heat,euler,batch")
    ap.add_argument("--tiles", type=int, default=10, help="tile heights per swept group")
    ap.add_argument("--copy-mib", type=int, default=256)
    a = ap.parse_args(argv)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if root not in sys.path:
        sys.path.insert(0, root)
    sw = importlib.import_module("1dsweptcuda_amd")
    groups = [g for g in a.groups.split(",") if g]
    n = 1 << 20
    for g in groups:
        if g == "copy":
            import torch

            x = torch.ones(a.copy_mib << 18, dtype=torch.float32, device="cuda")
            y = torch.empty_like(x)
            torch.cuda.synchronize()
            torch.neg(x, out=y)
            torch.cuda.synchronize()
            del x, y
            _say(f"copy: 1 launch over {a.copy_mib} MiB")
        elif g == "ks":
            k = _swept(sw, sw.KS, n, 1024, a.tiles, sw.ks_noisy_ic(n, 1), 0.005)
            _say(f"ks: {k} launches")
        elif g == "heat":
            x = np.arange(n) * 0.001
            ic = 100.0 * x * (1.0 - x / (0.001 * (n - 1)))
            k = _swept(sw, sw.HEAT, n, 512, a.tiles, ic, 1e-3)
            _say(f"heat: {k} launches")
        elif g == "euler":
            q = np.zeros((n, 3))
            q[: n // 2] = (1.0, 0.0, 1.0 / 0.4)
            q[n // 2:] = (0.125, 0.0, 0.1 / 0.4)
            k = _swept(sw, sw.EULER, n, 512, a.tiles, q.ravel(), 0.1 / (n - 2))
            _say(f"euler: {k} launches")
        elif g == "batch":
            wl = importlib.import_module("1dsweptcuda_amd.workloads")
            at = importlib.import_module("1dsweptcuda_amd.autotune")
            eq, nx, dt, dx = wl.batch_problems(4096, seed=2024)
            packed = np.concatenate(wl.batch_ics(eq, nx, dx))
            table = at.load()
            tps = at.tile_points_for(nx, table) if table else None
            _say("batch: initial conditions ready")
            for prec in (sw.F64, sw.F32):
                with sw.Batch(eq, nx, dt, dx, precision=prec, tile_points=tps, waves=at.load_waves()) as b:
                    b.set_state(packed)
                    b.advance(5000)
                    b.get_state()
                _say(f"batch {'f64' if prec == sw.F64 else 'f32'}: 1 launch, 5000 steps")
        else:
            raise SystemExit(f"unknown group {g!r}")


if __name__ == "__main__":
    main()
