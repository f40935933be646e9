"""Benchmark: point-updates/s of the 1D swept-rule solver on one MI355X.

Workload (default): Kuramoto-Sivashinsky, FP64, nX = 2^20, dx = 0.5,
dt = 0.005, IC 2cos(19 pi x/128) + N(0, 1e-3) noise (seed 1), shared swept
algorithm.  A "step" is one time step of the whole domain (nX point-updates).
The timed region is --steps steps rounded up to whole swept tile launches (a
diamond advances W/8 = 128 steps at node width 1024; at least 160 launches),
with the state already in HBM, bracketed by device synchronisation; "steps"
in the output is the count actually timed.  Warm-up is rounded the same way.
The classic algorithm (one launch per RK stage) is timed on the same input
for the swept/classic ratio.

The roofline entry is for the dominant kernel (the swept diamond): its
average duration comes from hipEvents recorded around every launch on the
plan's stream during a second pass over the same steps; achieved = algorithmic
FP64 flops per launch (40 flop per KS point-update x nX x W/8 steps) / that
duration.

The CPU baseline is the C oracle (the same KS stages) on a bounded sample of
the same workload (same nX, fewer steps), timed on this host with OpenMP over
the threads this job may use (at most 16) and with one thread.  The same leg
states the headline's error: the swept plan and the oracle advance the same
initial condition by the same whole number of tile launches and are compared.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

KS_FLOP_PER_PU = 40          # stutterStep 20 + finalStep 20 (KS1D_SweptShared.cu:159-171)
KS_SWEPT_BYTES_PER_PU = 256  # per tile launch: (u, u_half) read + written, 32 B / (W/8 steps) * W
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per whole-diamond launch (nX, node width) -> bytes, from rocprofv3
# FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by its reading for a copy
# kernel of known size (profiles/ks_swept_diamond_nx1048576_hbm_bytes.md).
DIAMOND_TRAFFIC_BYTES = {(1 << 20, 512): 33.96e6, (1 << 20, 1024): 34.77e6}
METRIC = "point-updates/s on 1 GPU at stated error vs FP64 oracle: swept vs classic/batch/NN"
# The reference's published figure for the same quantity: KS FP64 shared swept at
# nX = 2^20, 105.408 us per time step on a Tesla K40c
# (ResultPlots/ConferencePaper/BestRuns.csv:14).
REFERENCE_KS_SWEPT_PU_PER_S = (1 << 20) / 105.408e-6
# Timed and warm-up steps are whole tile launches (a swept advance of fewer
# steps than one tile runs classic steps instead), at least this many of them:
# 160 launches of the 2^20-point diamond are ~25 ms of GPU time.
MIN_TIMED_LAUNCHES = 160
MIN_WARM_LAUNCHES = 160


_T0 = time.perf_counter()


def _progress(section):
    """One line per finished section on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {section} done at {time.perf_counter() - _T0:.1f} s", file=sys.stderr, flush=True)


def _whole_launches(per, requested, minimum):
    """`requested` steps rounded up to whole tile launches of `per` steps, and
    at least `minimum` launches (per <= 1: the plan has no tiles)."""
    if per <= 1:
        return max(requested, minimum)
    return max(minimum, -(-requested // per)) * per


def _warm(plan, steps, minimum=40):
    """Untimed steps before a timed run: at least `steps`, rounded up to whole
    tile launches, and at least `minimum` launches, so the plan's one-time
    set-up (the launch graphs it captures on its first long advance) stays out
    of the timed region."""
    plan.advance(_whole_launches(plan.info()["tile_steps"], steps, minimum))


def _timed(plan, steps):
    plan.synchronize()
    t0 = time.perf_counter()
    plan.advance(steps)
    plan.synchronize()
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000,
                    help="timed time steps; rounded up to whole tile launches, at least %d" % MIN_TIMED_LAUNCHES)
    ap.add_argument("--warmup", type=int, default=20000,
                    help="untimed time steps; rounded up the same way, at least %d launches" % MIN_WARM_LAUNCHES)
    ap.add_argument("--nx", type=int, default=1 << 20)
    ap.add_argument("--node-width", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip every extra section (same as --extras none)")
    ap.add_argument("--extras", default="trapezoid,equations,sizes,batched,surrogate",
                    help="comma-separated extra sections to run, or 'none'")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one plain copy kernel over this many MiB (reads and writes "
                         "that many bytes), the known-size reference for HBM byte counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the solver runs on a single device, no collectives)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration_mib:
        _copy_calibration(args.copy_calibration_mib)

    sw = importlib.import_module("1dsweptcuda_amd")
    n, dt = args.nx, 0.005
    u0 = sw.ks_noisy_ic(n, args.seed)  # the field tests/test_gpu_fullsize.py checks bit for bit

    # ---------------------------------------------------------- swept (alg 1)
    plan = sw.Plan(sw.KS, n, dt, algorithm=sw.SWEPT, tpb=args.node_width, node_width=args.node_width)
    info = plan.info()
    H, W = info["tile_steps"], info["node_width"]
    if H <= 1:
        raise SystemExit(f"bench: nX={n} with node width {args.node_width} gives no swept tiles")
    steps = _whole_launches(H, args.steps, MIN_TIMED_LAUNCHES)
    warm = _whole_launches(H, args.warmup, MIN_WARM_LAUNCHES)
    plan.set_state(u0)
    plan.advance(warm)
    secs = _timed(plan, steps)
    value = n * steps / secs

    # per-kernel durations over the same number of steps: hipEvents recorded
    # around every launch on the plan's stream
    plan.set_profiling(True)
    plan.advance(steps)
    plan.synchronize()
    dn, dms = plan.kernel_stats(sw.K_DIAMOND)
    sn, sms = plan.kernel_stats(sw.K_SPLIT)
    plan.set_profiling(False)
    plan.close()
    launches = dn + sn
    if launches == 0 or dms + sms <= 0.0:
        raise SystemExit(f"bench: the profiled pass of {steps} steps recorded no diamond launch "
                         f"(tile height {H} steps)")
    avg_ms = (dms + sms) / launches
    pu_per_launch = n * H
    flops = KS_FLOP_PER_PU * pu_per_launch
    achieved_tf = flops / (avg_ms * 1e-3) / 1e12
    _progress("swept")

    # ------------------------------------------------------- classic (alg 0)
    cplan = sw.Plan(sw.KS, n, dt, algorithm=sw.CLASSIC, tpb=256)
    cplan.set_state(u0)
    cplan.advance(512)  # includes the capture of the 64-step launch graph
    csteps = max(steps // 4, 4096)
    csecs = _timed(cplan, csteps)
    classic_value = n * csteps / csecs
    cplan.close()
    # HBM roofline of the classic stages over the wall time of the timed steps
    # (both stage launches and anything between them): predictor reads u and
    # writes u_half (16 B), corrector reads u_half and u and writes u (24 B).
    classic_gbs = 40 * n * csteps / csecs / 1e9
    _progress("classic")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "ms_per_step": secs / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / REFERENCE_KS_SWEPT_PU_PER_S,
        "baseline": {"value": REFERENCE_KS_SWEPT_PU_PER_S, "unit": "point-updates/s",
                     "source": "reference KS FP64 shared swept, nX=2^20, 105.408 us/step on a Tesla K40c "
                               "(ResultPlots/ConferencePaper/BestRuns.csv:14)"},
        "dtype": "f64",
        "data": "synthetic: IC 2cos(19 pi x/128) + N(0,1e-3) drawn by sw1d_initial_condition "
                "(splitmix64, seed %d)" % args.seed,
        "config": {"workload": "Kuramoto-Sivashinsky 1D, RK2 + central differences, FP64, periodic, "
                               "shared swept (alg 1)",
                   "nx": n, "dx": 0.5, "dt": dt, "algorithm": "shared swept (alg 1)",
                   "node_width": W, "steps_per_tile_launch": H},
        "classic": {"value": classic_value, "steps": csteps, "ms_per_step": csecs / csteps * 1e3,
                    "roofline": {"bound": "hbm", "achieved": classic_gbs, "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": classic_gbs / HBM_PEAK_GBS,
                                 "algorithmic_bytes_per_pu": 40}},
        "swept_over_classic": value / classic_value,
        "roofline": {"bound": "fp64_valu", "compute": "FP64 vector ALU (no MFMA); peak = MI355X FP64 vector",
                     "kernel": "k_ks_swept<double,%d,DIAMOND>" % (W // 64),
                     "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tf / FP64_PEAK_TFLOPS,
                     "traffic": DIAMOND_TRAFFIC_BYTES.get((n, W)),
                     "avg_launch_us": avg_ms * 1e3, "launches": launches,
                     "wall_us_per_tile_launch": secs / (steps // H) * 1e6,
                     "flop_per_launch": flops,
                     "algorithmic_bytes_per_launch": 32 * n},
    }
    extras = set() if args.no_extras else set(args.extras.split(",")) - {"none", ""}
    sections = (("trapezoid", lambda: _trapezoid(sw, n, dt, u0, args)),
                ("equations", lambda: _equations(sw, n)),
                ("sizes", lambda: _sizes(sw, args.seed)),
                ("batched", lambda: _batched(sw)),
                ("surrogate", _surrogate))
    for name, run in sections:
        if name in extras:
            out[name] = run()
            _progress(name)
    if not args.no_cpu:
        out["cpu_baseline"] = _cpu_baseline(n, dt, u0, args.cpu_seconds)
        _progress("cpu_baseline")
        out["error_vs_fp64_oracle"] = _oracle_check(sw, n, dt, u0, W, out["cpu_baseline"]["check_steps"])
        _progress("oracle check")
    print(json.dumps(out))


def _copy_calibration(mib):
    """One elementwise kernel that reads `mib` MiB and writes `mib` MiB
    (torch.neg over float32), for scaling FETCH_SIZE / WRITE_SIZE readings."""
    import torch
    a = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.neg(a, out=b)
    torch.cuda.synchronize()
    del a, b


def _equations(sw, n):
    """Heat (the reference's parabola, insulated ends, Fo = 0.084) and Euler
    (Sod, dt/dx = 0.1) at the same nX, FP64: classic, shared swept, hybrid
    CPU+GPU swept (the host computes the boundary node's diamond) and
    trapezoid point-updates/s (every engine gives the oracle's bits)."""
    res = {}
    x = np.arange(n) * 0.001
    lx = 0.001 * (n - 1)
    heat_ic = 100.0 * x * (1.0 - x / lx)
    q = np.zeros((n, 3))
    q[: n // 2] = (1.0, 0.0, 1.0 / 0.4)
    q[n // 2:] = (0.125, 0.0, 0.1 / 0.4)
    dt_eu = 0.1 / (n - 2)
    cases = {
        "heat": (sw.HEAT, 1e-3, heat_ic, 20000, ((sw.CLASSIC, 256, 0), (sw.SWEPT, 512, 0), (sw.ALT, 512, 0),
                                                 (sw.TRAPEZOID, 256, 1))),
        "euler": (sw.EULER, dt_eu, q.ravel(), 2000, ((sw.CLASSIC, 256, 0), (sw.SWEPT, 512, 0), (sw.ALT, 512, 0),
                                                     (sw.TRAPEZOID, 256, 1))),
    }
    for name, (eq, dt, ic, steps, algs) in cases.items():
        row = {}
        for alg, W, depth in algs:
            p = sw.Plan(eq, n, dt, algorithm=alg, tpb=W, node_width=W, tile_depth=depth)
            p.set_state(ic)
            k = steps if alg != sw.CLASSIC else max(steps // 5, 100)
            _warm(p, k // 10)
            secs = _timed(p, k)
            p.close()
            row[{sw.CLASSIC: "classic", sw.SWEPT: "swept", sw.ALT: "hybrid", sw.TRAPEZOID: "trapezoid"}[alg]] = \
                n * k / secs
        row["swept_over_classic"] = row["swept"] / row["classic"]
        res[name] = row
    res["unit"] = "point-updates/s"
    return res


def _sizes(sw, seed, steps=20000):
    """The KS grid of the reference's performance sweeps, FP64: nX = 2^14 ..
    2^20 with swept node widths 64 .. 1024 (the reference's tpb), the
    classic algorithm and the register variant (alg 2), 20000 steps each;
    Euler (Sod) classic and swept over the same sizes; and KS FP32 at 2^20."""
    res = {}
    for lg in (14, 16, 18, 20):
        n = 1 << lg
        u0 = sw.ks_noisy_ic(n, seed)
        row = {"swept": {}}
        for alg, name, widths in ((sw.CLASSIC, "classic", (256,)), (sw.ALT, "register", (128,)),
                                  (sw.SWEPT, "swept", (64, 128, 256, 512, 1024))):
            for W in widths:
                p = sw.Plan(sw.KS, n, 0.005, algorithm=alg, tpb=W, node_width=W if alg == sw.SWEPT else 0)
                p.set_state(u0)
                _warm(p, steps // 10)
                secs = _timed(p, steps)
                p.close()
                if alg == sw.SWEPT:
                    row["swept"][str(W)] = n * steps / secs
                else:
                    row[name] = n * steps / secs
        best = max(row["swept"], key=row["swept"].get)
        row["best_swept"] = {"node_width": int(best), "value": row["swept"][best]}
        row["best_swept_over_classic"] = row["swept"][best] / row["classic"]
        res["2^%d" % lg] = row
        _progress(f"sizes KS 2^{lg}")
    # Euler (Sod, dt/dx = 0.1, FP64) over the same sizes: classic against the
    # swept node widths 64 .. 512 (Euler's range; narrow nodes give small
    # sizes more waves and less serial work per wave), 2000 steps each
    eu = {}
    for lg in (14, 16, 18, 20):
        n = 1 << lg
        q = np.zeros((n, 3))
        q[: n // 2] = (1.0, 0.0, 1.0 / 0.4)
        q[n // 2:] = (0.125, 0.0, 0.1 / 0.4)
        dt_eu = 0.1 / (n - 2)
        row = {"swept": {}}
        for alg, W in ((sw.CLASSIC, 256), (sw.SWEPT, 64), (sw.SWEPT, 128), (sw.SWEPT, 256), (sw.SWEPT, 512)):
            p = sw.Plan(sw.EULER, n, dt_eu, algorithm=alg, tpb=W, node_width=W if alg == sw.SWEPT else 0)
            W = p.info()["node_width"] if alg == sw.SWEPT else W
            p.set_state(q.ravel())
            _warm(p, 200)
            secs = _timed(p, 2000)
            p.close()
            if alg == sw.SWEPT:
                row["swept"][str(W)] = n * 2000 / secs
            else:
                row["classic"] = n * 2000 / secs
        best = max(row["swept"], key=row["swept"].get)
        row["best_swept"] = {"node_width": int(best), "value": row["swept"][best]}
        row["best_swept_over_classic"] = row["swept"][best] / row["classic"]
        eu["2^%d" % lg] = row
    res["euler_f64"] = eu
    # FP32 at the largest size: classic and the swept node widths that can win there
    n = 1 << 20
    u0 = sw.ks_noisy_ic(n, seed, precision=sw.F32)
    row = {}
    for alg, W, key in ((sw.CLASSIC, 256, "classic"), (sw.SWEPT, 512, "swept_512"), (sw.SWEPT, 1024, "swept_1024")):
        p = sw.Plan(sw.KS, n, 0.005, precision=sw.F32, algorithm=alg, tpb=W, node_width=W if alg == sw.SWEPT else 0)
        p.set_state(u0)
        _warm(p, steps // 10)
        secs = _timed(p, steps)
        p.close()
        row[key] = n * steps / secs
    res["f32_2^20"] = row
    res["unit"] = "point-updates/s"
    return res


def _trapezoid(sw, n, dt, u0, args):
    """Same workload on the overlapped trapezoid engine (algorithm 3), timed
    over the same requested steps rounded up to whole tile launches."""
    res = {}
    for W, depth in ((512, 2), (1024, 2)):
        p = sw.Plan(sw.KS, n, dt, algorithm=sw.TRAPEZOID, node_width=W, tile_depth=depth)
        H = p.info()["tile_steps"]
        steps = _whole_launches(H, args.steps, MIN_TIMED_LAUNCHES)
        p.set_state(u0)
        _warm(p, args.warmup, MIN_WARM_LAUNCHES)
        secs = _timed(p, steps)
        p.close()
        res["W%d_depth%d" % (W, depth)] = n * steps / secs
    best = max(res, key=res.get)
    return {"value": res[best], "config": best, "all": res, "unit": "point-updates/s"}


def _batched(sw, count=4096, steps=5000, loop_sample=512):
    """4096 uneven periodic heat/KS problems, 5000 steps, one persistent launch
    per precision, against a per-problem launch loop (trapezoid plans)."""
    wl = importlib.import_module("1dsweptcuda_amd.workloads")
    at = importlib.import_module("1dsweptcuda_amd.autotune")
    eq, nx, dt, dx = wl.batch_problems(count, seed=2024)
    ics = wl.batch_ics(eq, nx, dx)
    table = at.load()
    tps = at.tile_points_for(nx, table) if table else None
    waves = at.load_waves()
    total = float(nx.sum())
    res = {"problems": count, "steps": steps, "total_points": int(total),
           "tile_table": {str(k): v for k, v in table.items()} or "size heuristic",
           "waves_per_workgroup": waves}
    packed = np.concatenate(ics)
    for prec, name in ((sw.F64, "f64"), (sw.F32, "f32")):
        with sw.Batch(eq, nx, dt, dx, precision=prec, tile_points=tps, waves=waves) as b:
            b.set_state(packed)
            b.advance(steps)  # captures the graph
            b.set_state(packed)
            t0 = time.perf_counter()
            device_ms = b.advance_timed(steps)  # returns after the launch completes
            secs = time.perf_counter() - t0
        # state resident in HBM before and after; the read-back is not timed
        res[name] = {"value": total * steps / secs, "seconds": secs, "device_ms": device_ms}
    # per-problem launch loop on an evenly spaced sample, FP64
    idx = np.arange(0, count, max(count // loop_sample, 1))
    plans = []
    for i in idx:
        p = sw.Plan(int(eq[i]), int(nx[i]), float(dt[i]), algorithm=sw.TRAPEZOID, node_width=256,
                    bc=sw.BC_PERIODIC, dx=float(dx[i]))
        p.set_state(ics[i])
        _warm(p, 0)  # captures the plan's launch graphs outside the timed loop
        p.set_state(ics[i])
        plans.append(p)
    for p in plans:
        p.synchronize()
    t0 = time.perf_counter()
    for p in plans:
        p.advance(steps)
    for p in plans:
        p.synchronize()
    secs = time.perf_counter() - t0
    for p in plans:
        p.close()
    loop = float(nx[idx].sum()) * steps / secs
    res["per_problem_loop_f64"] = {"value": loop, "sample_problems": int(idx.size)}
    res["speedup_vs_loop_f64"] = res["f64"]["value"] / loop
    return res


def _surrogate():
    """Learned KS surrogate (nX=1024, K=20): trained on the GPU from 512 solver
    trajectories of 10000 steps, then a judged rollout (every 10th network
    step checked against the plain solver, tolerance 1e-2 relative L2) on 256
    unseen trajectories; throughput in point-update equivalents (problems x
    nX x K per network step)."""
    sur = importlib.import_module("1dsweptcuda_amd.surrogate")
    t0 = time.perf_counter()
    nn, rep = sur.train(ntraj=512, steps=10000, iters=3000, log=lambda *_: None)
    ev = sur.evaluate(nn, ntraj=256)
    nn.close()
    return {"value": ev["surrogate_pu_per_s"], "unit": "point-update equivalents/s",
            "speedup_vs_plain_batch": ev["speedup_vs_plain_batch"],
            "plain_batch_same_problems": ev["plain_batch_pu_per_s"],
            "note": "like for like: the plain batched solver on the same 256 problems and horizon; "
                    "the network costs ~8.2k flop per point-update equivalent against KS's 40",
            "judged_rollout": ev["judged_rollout"], "tolerance_per_judged_step": 1e-2,
            "one_step_rel_l2": ev["one_step_rel_l2"], "rollout_rel_l2": ev["rollout_rel_l2"],
            "held_out_rel_err_du": rep["held_out_rel_err_du"], "train_seconds": rep["train_seconds"],
            "dataset_samples": rep["samples"],
            "unjudged_value": ev["surrogate_unjudged_pu_per_s"],
            "plain_single_swept_nx1024": ev["plain_single_swept_pu_per_s"],
            "speedup_vs_single_swept": ev["speedup_vs_single_swept"],
            "seconds": time.perf_counter() - t0}


def _cpu_baseline(n, dt, u0, budget_s, repeats=5):
    """The C oracle (the same KS stages, FP64) on this host: OpenMP over the
    host threads this job may use (at most 16), and one thread.  For each,
    the step count is set so that one run takes about budget_s / (2 repeats)
    seconds (at least 1 s), and the median rate of `repeats` runs is kept;
    only the oracle call is timed (the state copy into it is ~1 ms).
    check_steps: the OpenMP runs' step count, which _oracle_check replays on
    the GPU."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    c = pyoracle.ks_consts(dt)

    def rate(threads, seconds):
        pyoracle.ks_run(u0, 2, c, threads)  # thread pool start-up, untimed
        steps, took = 2, 0.0
        while took < 0.1:  # calibrate on a run of at least 0.1 s
            steps *= 2
            t0 = time.perf_counter()
            pyoracle.ks_run(u0, steps, c, threads)
            took = time.perf_counter() - t0
        steps = max(2, int(seconds * steps / took))
        rates = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            pyoracle.ks_run(u0, steps, c, threads)
            rates.append(n * steps / (time.perf_counter() - t0))
        return float(np.median(rates)), steps

    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, avail))
    per_run = max(1.0, budget_s / (2 * repeats))
    omp, omp_steps = rate(threads, per_run)
    one, one_steps = rate(1, per_run)
    return {"value": omp, "unit": "point-updates/s", "cores": threads, "kind": "port",
            "sample": f"KS FP64 nX={n}, {omp_steps} steps per run, median of {repeats} runs, "
                      f"C oracle (oracle/swept1d_oracle.c) with OpenMP over {threads} threads",
            "serial": {"value": one, "cores": 1,
                       "sample": f"same, {one_steps} steps per run, median of {repeats}, one thread"},
            "host": {"cpu": _cpu_model(), "cpus_visible": os.cpu_count(), "cpus_allowed": avail},
            "check_steps": omp_steps}


def _oracle_check(sw, n, dt, u0, W, steps):
    """The stated error of the headline (part of the CPU-baseline leg): the
    swept plan of the timed configuration advances the same initial condition
    by `steps` steps (the CPU baseline's run length, rounded up to whole tile
    launches), and the OpenMP oracle's state after the same steps is the
    reference it is compared with."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    with sw.Plan(sw.KS, n, dt, algorithm=sw.SWEPT, tpb=W, node_width=W) as p:
        H = p.info()["tile_steps"]
        steps = _whole_launches(H, steps, 2)
        p.set_state(u0)
        p.advance(steps)
        got = p.get_state()
    want = pyoracle.ks_run(u0, steps, pyoracle.ks_consts(dt), max(1, min(16, avail)))
    diff = np.abs(got - want)
    scale = float(np.max(np.abs(want)))
    return {"steps": steps, "diamond_launches": max(steps // H - 1, 0), "bitwise_equal": bool(np.array_equal(got, want)),
            "max_abs_diff": float(diff.max()), "max_rel_diff": float(diff.max() / scale) if scale else 0.0,
            "tolerance_rel": 1e-12}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
