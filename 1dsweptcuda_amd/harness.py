"""Python 3 counterpart of the reference's run-and-analyse tools.

The reference drives its programs and reads their files with Python 2 helpers:
`runCUDA` and `Perform` (pyAnalysisTools/main_help.py:14-66), `Solved`
(pyAnalysisTools/result_help.py:24-52), `gatherTest` and `QualityRuns`
(pyAnalysisTools/analysis_help.py:31-66), and a test procedure with exact
solutions and a consistency run (SweptSource/Testing/testProcedure.py:51-237).
This module keeps their names, arguments and results for the programs of this
build (same argv, same .dat / .txt files), without plotting, restates the
timing sweep of SweptSource/runTest.sh (timing_sweep), and adds a reader for
the programs' full-precision snapshots (--snapshot=FILE).

    import importlib; h = importlib.import_module("1dsweptcuda_amd.harness")
    h.runCUDA(h.program("KS"), 1024, 64, 1e-3, 20, 6, 1, "ks.dat")
    s = h.Solved("ks.dat"); rows = s.stripInitial()   # {"Velocity": {t: values}}
"""
from __future__ import annotations

import collections
import os
import os.path as op
import shlex
import struct
import subprocess as sp

import numpy as np

BINDIR = op.join(op.dirname(op.abspath(__file__)), "bin")
GAM = 1.4
HEAT_ALPHA = 8.418e-5  # Heat1D_SweptShared.cu:56


def program(problem: str, precision: str = "Double") -> str:
    """Path of a built program, e.g. program("KS") -> .../bin/KSDoubleOut."""
    return op.join(BINDIR, problem + precision + "Out")


def makeList(v):
    return v if isinstance(v, list) else [v]


def runCUDA(Prog, divisions, threadsPerBlock, timeStep, finishTime, frequency, decomp,
            varfile="temp.dat", timefile="", options=(), quiet=False):
    """Run Prog once per (tpb, divisions) pair, as main_help.py:45-66 does;
    `options` are extra --key=value arguments.  Returns the exit codes."""
    codes = []
    for tpb in makeList(threadsPerBlock):
        for dvs in makeList(divisions):
            if not quiet:
                print("---------------------")
                print("Algorithm #divs #tpb dt endTime")
                print(decomp, dvs, tpb, timeStep, finishTime)
            cmd = "{0} {1} {2} {3} {4} {5} {6} {7} {8}".format(
                shlex.quote(Prog), dvs, tpb, timeStep, finishTime, frequency, decomp,
                shlex.quote(varfile), shlex.quote(timefile) if timefile else "")
            args = shlex.split(cmd) + list(options)
            r = sp.run(args, stdout=sp.DEVNULL if quiet else None)
            codes.append(r.returncode)
    return codes


class Perform:
    """Timing table (main_help.py:14-27): a header line, then rows of
    nX, tpb, microseconds per timestep; `dataMatrix` is pivoted to
    rows = nX, columns = tpb."""

    def __init__(self, dataset):
        import pandas as pd

        self.textfile = op.abspath(dataset)
        self.datafilename = op.splitext(op.basename(dataset))[0]
        self.fileLocation = op.dirname(dataset)
        raw = pd.read_table(dataset, sep=r"\s+")
        self.headers = [h.replace("_", " ") for h in raw.columns.values.tolist()]
        raw.columns = self.headers
        self.dataMatrix = raw.pivot_table(index=self.headers[0], columns=self.headers[1],
                                          values=self.headers[2], aggfunc="last")


TIMING_HEADER = "Num_Spatial_Points\tThreads_per_Block\tTime_per_timestep_(us)\n"


def write_timing_header(path):
    """The header the reference's sweeps write before the programs append rows
    (sweptPerformanceTest.py:232, quickperform.py:97)."""
    with open(path, "w") as f:
        f.write(TIMING_HEADER)


class Solved:
    """A .dat solution file (result_help.py:24-52): header `length nX dx`, then
    rows `Name t v0 v1 ...`."""

    def __init__(self, vFile):
        with open(vFile) as f:
            lines = [l for l in f.read().split("\n") if l.strip()]
        head = lines[0].split()
        self.datafilename = op.splitext(op.basename(vFile))[0]
        self.numpts = int(head[1])
        self.xGrid = np.linspace(0, float(head[0]), self.numpts)
        rows = [l.split() for l in lines[1:]]
        self.varNames = np.array([r[0] for r in rows])
        self.tFinal = np.around(np.array([float(r[1]) for r in rows]), decimals=7)
        self.vals = np.array([[float(v) for v in r[2:]] for r in rows])
        self.utFinal = np.unique(self.tFinal)
        self.plotTitles = np.unique(self.varNames)
        self.plotname = self.datafilename.split("_")[0]
        self.subpl = "Euler" in self.plotname

    def stripInitial(self):
        """{variable: {t: values}} for every row except t = 0."""
        stripped = collections.defaultdict(dict)
        for i, t in enumerate(self.tFinal):
            if t == 0:
                continue
            stripped[self.varNames[i]][t] = self.vals[i, :]
        return stripped


def gatherTest(resultPath):
    """All timing files named Problem_Precision_Algorithm_x.txt under resultPath
    in one frame indexed by (Problem, Precision, Algorithm, nX)
    (analysis_help.py:31-56)."""
    import pandas as pd

    names = ["Problem", "Precision", "Algorithm", "Num Spatial Points"]
    files = sorted(op.join(resultPath, f) for f in os.listdir(resultPath) if f.count("_") == 3)
    frames = []
    for f in files:
        m = Perform(f).dataMatrix
        key = op.basename(f).split("_")
        idx = [tuple(key[:3] + [i]) for i in m.index.get_level_values(0)]
        m.index = pd.MultiIndex.from_tuples(idx, names=names)
        frames.append(m)
    return pd.concat(frames)


# SweptSource/runTest.sh:8-38: the reference's timing sweep, 5e4 time steps of
# each problem at its time step, output frequency 2 tf (no intermediate rows).
SWEEP_PROBLEMS = (("Heat", 0.001), ("KS", 0.005), ("Euler", 1.0e-7))
SWEEP_SCHEMES = ("Classic", "Shared", "Alternative")


def timing_sweep(resultPath, nts=5e4, divisions=None, threadsPerBlock=None, precision="Double",
                 problems=None, progress=print):
    """The reference's timing sweep (SweptSource/runTest.sh:8-38): for every
    problem and scheme (algorithm 0, 1, 2), every nX = 2^11 .. 2^20 and
    tpb = 2^6 .. 2^10, `nts` time steps with output frequency 2 tf, each run
    appending its timing line to resultPath/<Problem>_<Precision>_<Scheme>_Timing.txt
    (after the header the reference's Python sweeps write).  gatherTest and
    QualityRuns read the files back.  Returns the files written."""
    os.makedirs(resultPath, exist_ok=True)
    divisions = divisions or [2 ** k for k in range(11, 21)]
    threadsPerBlock = threadsPerBlock or [2 ** k for k in range(6, 11)]
    files = []
    for prob, dt in SWEEP_PROBLEMS:
        if problems and prob not in problems:
            continue
        tf = dt * nts
        for alg, scheme in enumerate(SWEEP_SCHEMES):
            path = op.join(resultPath, "%s_%s_%s_Timing.txt" % (prob, precision, scheme))
            write_timing_header(path)
            if progress:
                progress("timing sweep: %s %s %s" % (prob, precision, scheme))
            runCUDA(program(prob, precision), divisions, threadsPerBlock, dt, tf, 2 * tf, alg,
                    op.join(resultPath, "temp.dat"), path, quiet=True)
            files.append(path)
    tmp = op.join(resultPath, "temp.dat")
    if op.exists(tmp):
        os.remove(tmp)
    return files


class QualityRuns:
    """Best time per row and how often each launch width wins
    (analysis_help.py:58-66)."""

    def __init__(self, dataMat):
        import pandas as pd

        self.bestRun = pd.DataFrame(dataMat.min(axis=1))
        best = dataMat.idxmin(axis=1)
        self.bestLaunch = best.value_counts()
        self.bestLaunch.index = pd.to_numeric(self.bestLaunch.index)
        self.bestLaunch = self.bestLaunch.sort_index()


# ------------------------------------------------------------- snapshots ---
SNAP_MAGIC = b"SW1DSNP1"
_SNAP_HEAD = struct.Struct("<8siiqqdd")


def read_snapshot(path):
    """(header, state) of a --snapshot=FILE: header has equation (0 heat, 1 KS,
    2 Euler), precision (0 double, 1 float), nx, components, time, dt; state is
    nx values (Euler: nx x 3, rows of rho, rho*u, E)."""
    with open(path, "rb") as f:
        raw = f.read()
    magic, eq, prec, nx, comps, t, dt = _SNAP_HEAD.unpack_from(raw)
    if magic != SNAP_MAGIC:
        raise ValueError(f"{path} is not a snapshot")
    data = np.frombuffer(raw, np.float64 if prec == 0 else np.float32, nx * comps, _SNAP_HEAD.size)
    head = {"equation": eq, "precision": prec, "nx": nx, "components": comps, "time": t, "dt": dt}
    return head, (data.reshape(nx, comps) if comps > 1 else data.copy())


# -------------------------------------------------------- exact solutions ---
def rmse(exact, sim):
    return float(np.sqrt(np.mean((np.asarray(exact) - np.asarray(sim)) ** 2)))


def heat_exact(t, divs, thing="Temperature", alpha=HEAT_ALPHA, terms=5000):
    """Series solution of the reference heat problem (testProcedure.py:63-82):
    insulated ends, parabolic start, L = divs*dx with dx = 0.001."""
    L = divs * 0.001
    xm = np.linspace(0, L, divs)
    n = np.arange(2, 2 * (terms + 1) + 1, 2)[:, None]
    h = np.sum(1.0 / n ** 2 * np.exp(-alpha * t * (n * np.pi / L) ** 2) * np.cos(n * xm[None, :] * np.pi / L),
               axis=0)
    return 50.0 * L / 3.0 - 400.0 * L / np.pi ** 2 * h


def sod_exact(x, t, x0=0.5, L=(1.0, 0.0, 1.0), R=(0.125, 0.0, 0.1), g=GAM):
    """Exact Riemann solution (Toro, Riemann Solvers and Numerical Methods for
    Fluid Dynamics, ch. 4): Newton on the pressure function, then sampling.
    Returns rows (density, velocity, energy e, pressure) at x, and the star
    state (p*, u*, rho*L, rho*R)."""
    rl, ul, pl = L
    rr, ur, pr = R
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)

    def fk(p, rk, pk, ak):
        if p > pk:
            A, B = 2 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            f = (p - pk) * np.sqrt(A / (p + B))
            df = np.sqrt(A / (B + p)) * (1 - (p - pk) / (2 * (B + p)))
        else:
            f = 2 * ak / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)
            df = 1 / (rk * ak) * (p / pk) ** (-(g + 1) / (2 * g))
        return f, df

    p = 0.5 * (pl + pr)
    for _ in range(60):
        f1, d1 = fk(p, rl, pl, al)
        f2, d2 = fk(p, rr, pr, ar)
        dp = (f1 + f2 + ur - ul) / (d1 + d2)
        p = max(p - dp, 1e-12)
        if abs(dp) < 1e-15 * p:
            break
    f1, _ = fk(p, rl, pl, al)
    f2, _ = fk(p, rr, pr, ar)
    us = 0.5 * (ul + ur) + 0.5 * (f2 - f1)
    rsl = rl * (p / pl) ** (1 / g)  # left rarefaction (p* < pL for Sod)
    rsr = rr * ((p / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pr + 1))
    x = np.asarray(x, np.float64)
    out = np.zeros((4, x.size))
    xi = (x - x0) / t
    shl, stl = ul - al, us - al * (p / pl) ** ((g - 1) / (2 * g))
    sr = ur + ar * np.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g))
    for i, s in enumerate(xi):
        if s < shl:
            r, u, pp = rl, ul, pl
        elif s < stl:
            u = 2 / (g + 1) * (al + (g - 1) / 2 * ul + s)
            a = 2 / (g + 1) * (al + (g - 1) / 2 * (ul - s))
            r = rl * (a / al) ** (2 / (g - 1))
            pp = pl * (a / al) ** (2 * g / (g - 1))
        elif s < us:
            r, u, pp = rsl, us, p
        elif s < sr:
            r, u, pp = rsr, us, p
        else:
            r, u, pp = rr, ur, pr
        out[:, i] = (r, u, pp / ((g - 1) * r), pp)
    return out, (p, us, rsl, rsr)


def euler_exact(t, divs, thing):
    """The Sod solution on the program's cell centres dx = 1/(divs-2)
    (testProcedure.py:84-96); thing is Density, Velocity, Energy or Pressure."""
    dx = 1.0 / float(divs - 2)
    # the divs-2 centres (j + 1/2) dx; a float-step arange can drop the last one
    x = (np.arange(divs - 2) + 0.5) * dx
    rows, _ = sod_exact(x, t)
    return rows[["density", "velocity", "energy", "pressure"].index(thing.lower())]


def ks_exact(t, divs, thing, fixture):
    """Row `thing` at time t of a KS reference solution file
    (testProcedure.py:98-99 reads KSDouble_Official.txt)."""
    return Solved(fixture).stripInitial()[thing][t]


# ------------------------------------------------------- test procedure ---
def consistency(problem, tf, dt=1e-5, div=4096, tpb=128, precision="Double", workdir=".", options=()):
    """testProcedure.py:106-122: run algorithms 1, 2 and 0 with one output at
    the end and collect (names, times, values) of each; the final rows of the
    first two are compared with the third's by RMSE."""
    exe = program(problem, precision)
    vfile = op.join(workdir, "temp.dat")
    collect = []
    for s in (1, 2, 0):
        codes = runCUDA(exe, div, tpb, dt, tf, tf * 2.0, s, vfile, options=options, quiet=True)
        if codes[0] != 0:
            raise RuntimeError(f"{exe} exited with {codes[0]}")
        a = Solved(vfile)
        collect.append((a.varNames, a.tFinal, a.vals))
        tf = a.tFinal[-1]
    return collect


def consistency_report(problem, tf, tol=1e-5, **kw):
    """{(algorithm, variable): rmse against classic} at the final time, and
    whether every one is within tol (testProcedure.py:209-237)."""
    rslt = consistency(problem, tf, **kw)
    tfin = rslt[-1][1][-1]
    col = []
    for names, times, vals in rslt:
        col.append({names[i]: vals[i] for i, y in enumerate(times) if y == tfin})
    out = {}
    for a, c in zip(("Swept", "Alternative"), col[:2]):
        for k in c:
            out[(a, k)] = rmse(c[k], col[2][k])
    return out, all(v <= tol for v in out.values())
