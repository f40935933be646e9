"""Training and evaluation driver of the learned KS surrogate.

Everything runs on the GPU through the C ABI: trajectories from the batched
solver (sw1d_batch_*), the training set built on the device from their
snapshots (sw1d_nn_dataset), Adam training with the BF16 MFMA forward and
backward kernels (sw1d_nn_fit), and the judged rollout (sw1d_surrogate_advance).

Workload: KS, nX = 1024, dx = 0.5, dt = 0.005, K = 20 solver steps per network
step; training trajectories start from sum_{k<=16} (a_k cos + b_k sin)(2 pi k x/L)/k,
a_k, b_k ~ N(0,1), seeded per trajectory.

    python 1dsweptcuda_amd/surrogate.py [--traj 512] [--steps 10000] [--iters 3000]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

NX, DX, DT, K = 1024, 0.5, 0.005, 20
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned", "ks_surrogate.npz")


def _pkg():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if root not in sys.path:
        sys.path.insert(0, root)
    return (importlib.import_module("1dsweptcuda_amd"),
            importlib.import_module("1dsweptcuda_amd.workloads"))


def ks_batch(sw, wl, ntraj, seed):
    """The plain batched solver on `ntraj` seeded trajectories, with the tuned
    waves per workgroup (the tile width of nX = 1024 is the default 512)."""
    at = importlib.import_module("1dsweptcuda_amd.autotune")
    b = sw.Batch(np.full(ntraj, sw.KS, np.int32), np.full(ntraj, NX, np.int64), np.full(ntraj, DT),
                 np.full(ntraj, DX), waves=at.load_waves())
    b.set_state(np.concatenate([wl.fourier_ic(NX, NX * DX, seed + i) for i in range(ntraj)]))
    return b


def train(ntraj=512, steps=10000, iters=3000, batch=64, lr=2e-3, seed=0, log=print):
    """Train on `ntraj` trajectories of `steps` solver steps; returns (nn, report)."""
    sw, wl = _pkg()
    rep = {"trajectories": ntraj, "steps": steps, "K": K, "nx": NX, "dt": DT, "dx": DX}
    nn = sw.Surrogate(NX, max(batch, 256), seed=seed)
    t0 = time.perf_counter()
    with ks_batch(sw, wl, ntraj, seed=1000 + seed) as b:
        nsamp, scales = nn.dataset(b, steps // K, K)
    rep["dataset_seconds"] = time.perf_counter() - t0
    rep["samples"], rep["scales"] = nsamp, scales
    # the last 10% of the snapshots (the latest times of every trajectory) are held out
    nfit = int(nsamp * 0.9) // ntraj * ntraj
    rep["held_out"] = nsamp - nfit
    t0 = time.perf_counter()
    done, hist = 0, []
    for frac, rate in ((0.5, lr), (0.3, lr / 4), (0.2, lr / 16)):
        n = int(iters * frac)
        loss = nn.fit(n, batch=batch, lr=rate, nfit=nfit, seed=seed + done)
        done += n
        err = nn.eval(nfit, min(nsamp - nfit, 4096))
        hist.append({"iters": done, "lr": rate, "loss": loss, "held_out_rel_err_du": err})
        log(json.dumps(hist[-1]))
    rep["train_seconds"] = time.perf_counter() - t0
    rep["history"] = hist
    rep["held_out_rel_err_du"] = hist[-1]["held_out_rel_err_du"]
    return nn, rep


def evaluate(nn, ntraj=256, warm=2000, network_steps=200, judge_every=10, tol=1e-2, seed=7):
    """Judged rollout on unseen trajectories against the plain solver: error,
    fallbacks, and throughput of both paths on the same problems."""
    sw, wl = _pkg()
    rep = {"trajectories": ntraj, "network_steps": network_steps, "judge_every": judge_every, "tol": tol}
    with ks_batch(sw, wl, ntraj, seed=5000 + seed) as b, ks_batch(sw, wl, ntraj, seed=5000 + seed) as ref:
        b.advance(warm)
        ref.advance(warm)
        u0 = b.get_state()
        # throughput: network steps without judging
        nn.advance(b, 2, K, judge_every=0)
        b.set_state(u0)
        t0 = time.perf_counter()
        st = nn.advance(b, network_steps, K, judge_every=judge_every, tol=tol)
        secs = time.perf_counter() - t0
        pu = float(ntraj) * NX * K * network_steps
        rep["judged_rollout"] = st
        rep["surrogate_pu_per_s"] = pu / secs
        judged = b.get_state().reshape(ntraj, NX)
        b.set_state(u0)
        t0 = time.perf_counter()
        nn.advance(b, network_steps, K, judge_every=0)
        secs_nj = time.perf_counter() - t0
        rep["surrogate_unjudged_pu_per_s"] = pu / secs_nj
        unjudged = b.get_state().reshape(ntraj, NX)
        # plain batched solver, same problems and horizon (state stays in HBM)
        ref.advance(K * network_steps)
        ref.set_state(u0)
        t0 = time.perf_counter()
        ref.advance_timed(K * network_steps)
        rep["plain_batch_pu_per_s"] = pu / (time.perf_counter() - t0)
        # accumulated error of the whole rollout against the plain solver's
        # trajectory from the same states, per problem (KS is chaotic: the
        # distance of two trajectories grows with the horizon whatever its start)
        plain_end = ref.get_state().reshape(ntraj, NX)
        norm = np.linalg.norm(plain_end, axis=1)
        rep["rollout_rel_l2"] = {
            "horizon_solver_steps": K * network_steps,
            "judged": {"median": float(np.median(np.linalg.norm(judged - plain_end, axis=1) / norm)),
                       "max": float(np.max(np.linalg.norm(judged - plain_end, axis=1) / norm))},
            "unjudged": {"median": float(np.median(np.linalg.norm(unjudged - plain_end, axis=1) / norm)),
                         "max": float(np.max(np.linalg.norm(unjudged - plain_end, axis=1) / norm))}}
        # error of the unjudged network after one network step, per trajectory
        ref.set_state(u0)
        ref.advance(K)
        plain = ref.get_state().reshape(ntraj, NX)
        b.set_state(u0)
        nn.advance(b, 1, K, judge_every=0)
        net = b.get_state().reshape(ntraj, NX)
        e = np.linalg.norm(net - plain, axis=1) / np.linalg.norm(plain, axis=1)
        rep["one_step_rel_l2"] = {"median": float(np.median(e)), "max": float(e.max())}
    # single plain plan at nX = 1024 (the reference's configuration)
    p = sw.Plan(sw.KS, NX, DT, algorithm=sw.SWEPT, tpb=128)
    p.set_state(u0[:NX])
    p.advance(K * 50)
    p.synchronize()
    t0 = time.perf_counter()
    p.advance(K * network_steps)
    p.synchronize()
    rep["plain_single_swept_pu_per_s"] = NX * K * network_steps / (time.perf_counter() - t0)
    p.close()
    rep["speedup_vs_plain_batch"] = rep["surrogate_pu_per_s"] / rep["plain_batch_pu_per_s"]
    rep["speedup_vs_single_swept"] = rep["surrogate_pu_per_s"] / rep["plain_single_swept_pu_per_s"]
    return rep


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--traj", type=int, default=512)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-traj", type=int, default=256)
    ap.add_argument("--save", default="")
    ap.add_argument("--report", default="")
    a = ap.parse_args(argv)
    nn, rep = train(a.traj, a.steps, a.iters, a.batch, a.lr)
    rep["evaluation"] = evaluate(nn, ntraj=a.eval_traj)
    if a.save:
        np.savez(a.save, params=nn.get_params(), scales=np.array(rep["scales"]))
    nn.close()
    out = json.dumps(rep)
    if a.report:
        with open(a.report, "w") as f:
            f.write(out + "\n")
    print(out)


if __name__ == "__main__":
    main()
