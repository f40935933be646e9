"""bench.py with a timed step count below one swept tile launch
(--steps 20 --warmup 5): the timed and warm-up steps are rounded up to whole
diamond launches, the profiled pass records diamonds, and the line carries
its roofline.  Runs bench.py as its own process on the GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_bench_short_step_count():
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "1", "--steps", "20", "--warmup", "5",
           "--no-cpu", "--extras", "none"]
    print("running:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    line = json.loads(r.stdout.strip().splitlines()[-1])
    H = line["config"]["steps_per_tile_launch"]
    assert H == line["config"]["node_width"] // 8 == 128
    assert line["steps"] % H == 0 and line["steps"] // H >= 160
    assert line["warmup"] % H == 0 and line["steps_requested"] == 20
    roof = line["roofline"]
    assert roof["bound"] == "fp64_valu" and roof["launches"] >= 1
    # a diamond launch advances nX points by H steps at 40 flop each
    assert roof["flop_per_launch"] == 40 * line["config"]["nx"] * H
    assert 0.0 < roof["frac"] < 1.0 and roof["avg_launch_us"] > 0
    assert line["value"] > line["classic"]["value"] > 0
    assert abs(line["ms_per_step"] * 1e-3 * line["value"] - line["config"]["nx"]) < 1e-6 * line["config"]["nx"]
