"""CPU emulation of the swept tiling used by the HIP kernels (test infrastructure).

Every point carries two values: heat the even- and odd-level rows, KS
(u, u_half), Euler the conserved triple and its predictor.  Stages alternate
between writing the second value from the first (odd s) and the first from the
second (even s).  The GPU tiles update, at stage s of a launch, exactly the points of each node
whose node-local position x satisfies |2x-(W-1)| <= thr(s), with
    down phase  thr = 2*R*s - 1          s = 1 .. W/(2R)
    up phase    thr = W - 1 - 2*R*m      m = 1 .. W/(2R) - 1
(R = stencil radius of one stage), nodes placed at offset 0 or n - W/2 on
alternate launches, up-triangle first and down-triangle last.  This module
replays that schedule on the CPU: every stage evaluates the oracle's classic
stage over the whole domain and keeps the result only at the active points.
Because all nodes of a launch step through the same stages together and a
node's halo is read only at the full-width stage (when the neighbour node has
not yet touched it), this is exactly what the kernels compute.  Comparing it
with the oracle's classic run checks the decomposition itself (active sets,
offsets, tile counts), independently of the GPU.
"""
from __future__ import annotations

import numpy as np

import pyoracle as po


def _mask(n, W, shifted, thr):
    base = (n - W // 2) if shifted else 0
    x = (np.arange(n) - base) % W
    return np.abs(2 * x - (W - 1)) <= thr


class Emu:
    def __init__(self, eq, n, dt, dtype=np.float64, bc="mirror", W=128, dx=None, alpha=8.418e-5):
        self.eq, self.n, self.W, self.dtype = eq, n, W, np.dtype(dtype)
        self.bc = bc
        if eq == "heat":
            self.R = 1
            self.fo, self.fo2 = po.heat_consts(dt, dtype, alpha=alpha, ds=dx or 0.001)
            self.H = W // 2
        elif eq == "ks":
            self.R = 2
            self.coef = po.ks_consts(dt, dtype, dx=dx or 0.5)
            self.H = W // 8
        else:
            self.R = 2
            self.k, self.bd = po.euler_consts(dt, n, dtype)
            self.H = W // 8

    # -- one stage over the whole domain -------------------------------------
    def _stage(self, st, a, b):
        """Candidate values of the component written by stage st."""
        if self.eq == "heat":
            return po.heat_step(a, self.fo, self.fo2, 0 if self.bc == "mirror" else 1)
        if self.eq == "ks":
            if st == 0:
                out = np.empty_like(a)
                po.ks_stage(a, out, self.coef, 0)
                return out
            out = a.copy()  # u
            po.ks_stage(b, out, self.coef, 1)
            return out
        if st == 0:
            out = np.empty_like(a)
            po.euler_stage(a, out, self.k, self.bd, 0)
            return out
        out = a.copy()
        po.euler_stage(b, out, self.k, self.bd, 1)
        return out

    def _tile(self, state, shifted, down, up):
        n, W, R = self.n, self.W, self.R
        Hs = W // (2 * R)
        seq = []
        if down:
            seq += [(2 * R * s - 1, (s + 1) % 2) for s in range(1, Hs + 1)]
        if up:
            seq += [(W - 1 - 2 * R * m, (m + 1) % 2) for m in range(1, Hs)]
        for thr, st in seq:
            mk = _mask(n, W, shifted, thr)
            if self.eq == "heat":
                # two rows: even levels in row 0, odd levels in row 1
                src, dst = (0, 1) if st == 0 else (1, 0)
                cand = self._stage(0, state[src], None)
                state[dst][mk] = cand[mk]
                continue
            if self.eq == "euler":
                mk3 = np.repeat(mk, 3)
            else:
                mk3 = mk
            u, uh = state
            if st == 0:
                cand = self._stage(0, u, None)
                uh[mk3] = cand[mk3]
            else:
                cand = self._stage(1, u, uh)
                u[mk3] = cand[mk3]

    def run(self, u0, nsteps):
        u = np.array(u0, dtype=self.dtype, copy=True)
        state = [u, u.copy()]
        tiles = nsteps // self.H
        rem = nsteps - tiles * self.H
        if tiles >= 1:
            shifted = False
            self._tile(state, shifted, False, True)
            for _ in range(tiles - 1):
                shifted = not shifted
                self._tile(state, shifted, True, True)
            shifted = not shifted
            self._tile(state, shifted, True, False)
        u = state[0]
        if rem:
            if self.eq == "heat":
                u = po.heat_run(u, rem, self.fo, self.fo2, 0 if self.bc == "mirror" else 1)
            elif self.eq == "ks":
                u = po.ks_run(u, rem, self.coef)
            else:
                u = po.euler_run(u, rem, self.k, self.bd)
        return u
