"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message and exit status 2 before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_cpu_baseline_fields():
    """The CPU baseline entry: the oracle timed with OpenMP and with one
    thread, median of the repeats, and the host it ran on."""
    sys.path.insert(0, ROOT)
    import bench

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    r = bench._cpu_baseline(4096, 0.005, pyoracle.ic_ks_noisy(4096, 1), 0.5, repeats=3)
    assert r["kind"] == "port" and r["unit"] == "point-updates/s"
    assert r["value"] > 0 and r["serial"]["value"] > 0 and r["serial"]["cores"] == 1
    assert 1 <= r["cores"] <= 16 and r["cores"] <= r["host"]["cpus_allowed"]
    assert "median of 3" in r["sample"] and r["host"]["cpu"]
    assert r["check_steps"] >= 2


def test_whole_launch_rounding():
    """Timed and warm-up steps are whole tile launches, at least the minimum:
    the driver's --steps 20 gives 160 launches of the 128-step diamond."""
    sys.path.insert(0, ROOT)
    import bench

    assert bench._whole_launches(128, 20, 160) == 160 * 128
    assert bench._whole_launches(128, 20001, 2) == 157 * 128
    assert bench._whole_launches(128, 128 * 200, 160) == 200 * 128
    assert bench._whole_launches(1, 20, 40) == 40
