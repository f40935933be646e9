"""Tile-size autotuner for the batched solver.

The reference picks its threads-per-block by brute force: every run of a tpb
sweep appends a timing line, and the analysis keeps the fastest configuration
per problem size (pyAnalysisTools/analysis_help.py:31-64, `QualityRuns`).  The
same idea here, on the device and per size class: problems are grouped by
floor(log2 nX); for each class a sample batch is timed with every candidate
tile width (output points per trapezoid tile), and the class gets the fastest.

`choose` is a pure function of the timing table, so a given set of timings
always gives the same choice: the minimum time per point-update wins, and
candidates within `tie` of the minimum resolve to the wider tile (fewer
redundant halo points).  Classes with no timing take the choice of the
nearest timed class below, else above.

A second pass tunes the waves per workgroup of the persistent launch (the
tiles a workgroup works on at once, the reference's "diamonds per block"):
the whole workload with the chosen tiles is timed at 2, 4 and 8 waves and
`choose_waves` keeps the fastest, the default 4 when it is within `tie`.

A third table serves single plans of the shared swept algorithm: the node
width (the reference's tpb, the points one diamond spans) per equation,
precision and size class, timed the same way and chosen by `choose`.
`Plan(..., node_width="auto")` takes its width from that table
(`node_width_for`); the programs keep the width the caller's tpb gives, as
the reference's do.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

CANDIDATES = (128, 256, 512, 1024)
WAVE_CANDIDATES = (2, 4, 8)
DEFAULT_WAVES = 4
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned", "batch_tiles.json")

NODE_CANDIDATES = (64, 128, 256, 512, 1024)
NODE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned", "node_widths.json")
EQ_NAMES = ("heat", "ks", "euler")
PREC_NAMES = ("f64", "f32")
# timed steps per node-width candidate: a few launches of the widest node
# (heat diamonds advance W/2 steps per launch, KS and Euler W/8)
NODE_STEPS = {0: 4096, 1: 1024, 2: 256}


def size_class(n: int) -> int:
    return int(math.floor(math.log2(int(n))))


def choose(timings: dict, tie: float = 0.02) -> dict:
    """timings: {class: {tile_points: seconds per point-update}} -> {class: tile_points}."""
    out = {}
    for cls in sorted(timings):
        row = {int(k): float(v) for k, v in timings[cls].items() if v is not None and v > 0}
        if not row:
            continue
        best = min(row.values())
        near = [tp for tp, v in row.items() if v <= best * (1.0 + tie)]
        out[int(cls)] = max(near)
    return out


def choose_waves(timings: dict, tie: float = 0.02) -> int:
    """timings: {waves: seconds per point-update} -> waves per workgroup."""
    row = {int(k): float(v) for k, v in timings.items() if v is not None and v > 0}
    if not row:
        return DEFAULT_WAVES
    best = min(row.values())
    if DEFAULT_WAVES in row and row[DEFAULT_WAVES] <= best * (1.0 + tie):
        return DEFAULT_WAVES
    return min(row, key=lambda w: (row[w], w))


def tile_points_for(nx, table: dict, default: int = 256) -> np.ndarray:
    """Per-problem tile width from a {class: tile_points} table."""
    classes = sorted(int(c) for c in table)
    res = np.empty(len(nx), np.int32)
    for i, n in enumerate(nx):
        c = size_class(n)
        if c in table:
            res[i] = table[c]
        elif classes:
            below = [k for k in classes if k <= c]
            res[i] = table[below[-1]] if below else table[classes[0]]
        else:
            res[i] = default
    return res


def tune_batch(sw, eq, nx, dt, dx, ics, precision: int = 0, steps: int = 400, sample: int = 2048,
               candidates=CANDIDATES, reps: int = 2, seed: int = 0):
    """Time every candidate per size class on the device; return (table, timings).

    Each class is timed as a batch of `sample` of its problems (drawn with
    repetition when the class has fewer), enough workgroups to fill every CU,
    so the timing is the class's throughput in a full launch rather than the
    latency of a few problems."""
    rng = np.random.default_rng(seed)
    eq, nx, dt, dx = map(np.asarray, (eq, nx, dt, dx))
    timings = {}
    for cls in sorted({size_class(n) for n in nx}):
        idx = np.flatnonzero([size_class(n) == cls for n in nx])
        idx = np.sort(rng.choice(idx, sample, replace=idx.size < sample))
        pts = float(nx[idx].sum()) * steps
        packed = np.concatenate([ics[i] for i in idx])
        row = {}
        for tp in candidates:
            if tp > 2 * int(nx[idx].min()):
                continue
            with sw.Batch(eq[idx], nx[idx], dt[idx], dx[idx], precision=precision,
                          tile_points=np.full(idx.size, tp, np.int32)) as b:
                b.set_state(packed)
                b.advance(steps)  # instantiate the graph, warm the caches
                ms = min(b.advance_timed(steps) for _ in range(reps))
            row[tp] = ms * 1e-3 / pts
        timings[cls] = row
    return choose(timings), timings


def tune_waves(sw, eq, nx, dt, dx, ics, table: dict, precision: int = 0, steps: int = 400, reps: int = 2,
               candidates=WAVE_CANDIDATES):
    """Time the whole workload, tiles from `table`, at each waves-per-workgroup
    candidate; return (choice, {waves: seconds per point-update})."""
    tps = tile_points_for(nx, table)
    packed = np.concatenate(ics)
    pts = float(np.sum(nx)) * steps
    timings = {}
    for w in candidates:
        with sw.Batch(eq, nx, dt, dx, precision=precision, tile_points=tps, waves=w) as b:
            b.set_state(packed)
            b.advance(steps)
            ms = min(b.advance_timed(steps) for _ in range(reps))
        timings[w] = ms * 1e-3 / pts
    return choose_waves(timings), timings


def save(table: dict, timings: dict, path: str = DEFAULT_TABLE, meta: dict | None = None,
         waves: int = DEFAULT_WAVES, wave_timings: dict | None = None) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"table": {str(k): v for k, v in table.items()},
                   "seconds_per_point_update": {str(k): {str(t): s for t, s in r.items()}
                                                for k, r in timings.items()},
                   "waves_per_workgroup": int(waves),
                   "seconds_per_point_update_by_waves": {str(w): s for w, s in (wave_timings or {}).items()},
                   "meta": meta or {}, "written": time.strftime("%Y-%m-%d %H:%M:%S")}, f, indent=1)


def load(path: str = DEFAULT_TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(k): int(v) for k, v in json.load(f)["table"].items()}


def load_waves(path: str = DEFAULT_TABLE) -> int:
    if not os.path.exists(path):
        return DEFAULT_WAVES
    with open(path) as f:
        return int(json.load(f).get("waves_per_workgroup", DEFAULT_WAVES))


def _plan_case(eq: int, n: int, rng):
    """A timing case of the equation at n points: the reference's initial
    condition and time step (heat Fo = 0.084 on the parabola, KS dt = 0.005,
    Sod at dt/dx = 0.1)."""
    if eq == 0:
        x = np.arange(n) * 0.001
        return 100.0 * x * (1.0 - x / (0.001 * (n - 1))), 1e-3
    if eq == 1:
        x = np.arange(n) * 0.5
        return 2.0 * np.cos(19.0 * np.pi * x / 128.0) + 1e-3 * rng.standard_normal(n), 0.005
    q = np.zeros((n, 3))
    q[: n // 2] = (1.0, 0.0, 1.0 / 0.4)
    q[n // 2:] = (0.125, 0.0, 0.1 / 0.4)
    return q.ravel(), 0.1 / (n - 2)


def tune_node_widths(sw, equations=(0, 1, 2), precisions=(0, 1), log2_sizes=(12, 14, 16, 18, 20),
                     candidates=NODE_CANDIDATES, reps: int = 2, seed: int = 0):
    """Time the shared swept plan at every candidate node width per equation,
    precision and size class; return (table, timings) keyed by
    (equation, precision) then size class.  A width the plan would clamp (an
    equation's range) or that does not split n into two nodes is skipped."""
    rng = np.random.default_rng(seed)
    table, timings = {}, {}
    for eq in equations:
        for prec in precisions:
            rows = {}
            for lg in log2_sizes:
                n = 1 << lg
                ic, dt = _plan_case(eq, n, rng)
                steps = NODE_STEPS[eq]
                row = {}
                for w in candidates:
                    if n < 2 * w:
                        continue
                    with sw.Plan(eq, n, dt, precision=prec, algorithm=sw.SWEPT, tpb=w, node_width=w) as p:
                        if p.info()["node_width"] != w:
                            continue
                        p.set_state(ic)
                        p.advance(steps // 4)
                        ms = min(p.advance_timed(steps) for _ in range(reps))
                    row[w] = ms * 1e-3 / (n * steps)
                rows[lg] = row
            timings[(eq, prec)] = rows
            table[(eq, prec)] = choose(rows)
    return table, timings


def save_node_widths(table: dict, timings: dict, path: str = NODE_TABLE, meta: dict | None = None) -> None:
    def nest(d, leaf):
        out = {}
        for (eq, prec), v in d.items():
            out.setdefault(EQ_NAMES[eq], {})[PREC_NAMES[prec]] = leaf(v)
        return out
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"table": nest(table, lambda r: {str(k): int(v) for k, v in r.items()}),
                   "seconds_per_point_update": nest(timings, lambda r: {str(k): {str(w): s for w, s in c.items()}
                                                                       for k, c in r.items()}),
                   "meta": meta or {}, "written": time.strftime("%Y-%m-%d %H:%M:%S")}, f, indent=1)


def load_node_widths(path: str = NODE_TABLE) -> dict:
    """{(equation, precision): {size class: node width}}; {} without a table."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        t = json.load(f)["table"]
    return {(EQ_NAMES.index(e), PREC_NAMES.index(pr)): {int(k): int(v) for k, v in r.items()}
            for e, byp in t.items() for pr, r in byp.items()}


def node_width_for(equation: int, n: int, precision: int = 0, table: dict | None = None) -> int:
    """The tuned node width of a single swept plan: the entry of n's size
    class, else of the nearest class below, else above; 0 (the caller's tpb)
    when the table has no entry for this equation and precision."""
    t = load_node_widths() if table is None else table
    row = t.get((int(equation), int(precision)))
    if not row:
        return 0
    c = size_class(n)
    if c in row:
        return row[c]
    classes = sorted(row)
    below = [k for k in classes if k <= c]
    return row[below[-1]] if below else row[classes[0]]


def main(argv=None):
    """Tune the batched workload's tile table and waves per workgroup on the
    current GPU, or with --plans the node widths of single swept plans.

    python 1dsweptcuda_amd/autotune.py [--out PATH] [--steps N] [--precision 0|1]
    python 1dsweptcuda_amd/autotune.py --plans [--out PATH]
    """
    import argparse
    import importlib
    import sys

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--plans", action="store_true", help="tune single-plan node widths instead")
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--count", type=int, default=4096)
    a = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    sw = importlib.import_module("1dsweptcuda_amd")
    if a.plans:
        table, timings = tune_node_widths(sw)
        save_node_widths(table, timings, a.out or NODE_TABLE,
                         meta={"algorithm": "shared swept (1)", "log2_sizes": [12, 14, 16, 18, 20],
                               "steps": {EQ_NAMES[k]: v for k, v in NODE_STEPS.items()}})
        print(json.dumps({"%s/%s" % (EQ_NAMES[e], PREC_NAMES[p]): r for (e, p), r in table.items()}))
        return
    a.out = a.out or DEFAULT_TABLE
    wl = importlib.import_module("1dsweptcuda_amd.workloads")
    eq, nx, dt, dx = wl.batch_problems(a.count, seed=2024)
    ics = wl.batch_ics(eq, nx, dx)
    table, timings = tune_batch(sw, eq, nx, dt, dx, ics, precision=a.precision, steps=a.steps)
    waves, wave_timings = tune_waves(sw, eq, nx, dt, dx, ics, table, precision=a.precision, steps=a.steps)
    save(table, timings, a.out, meta={"precision": a.precision, "steps": a.steps,
                                      "workload": "batch_problems(%d, seed=2024)" % a.count},
         waves=waves, wave_timings=wave_timings)
    print(json.dumps({"table": table, "waves_per_workgroup": waves}))


if __name__ == "__main__":
    main()
