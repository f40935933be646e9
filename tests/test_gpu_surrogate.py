"""Learned KS surrogate on the GPU: the BF16 MFMA forward and backward kernels
against the CPU oracle (oracle/nn_oracle.py, whose backward is pinned by
finite differences), GPU training, and the judged fallback."""
import numpy as np
import pytest

import nn_oracle as no

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module")
def wl():
    import importlib

    return importlib.import_module("1dsweptcuda_amd.workloads")


def test_mfma_operand_layout(sw):
    print("mfma layout check", flush=True)
    assert sw.Surrogate.device_check()


def _params(seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(no.NPARAM) * 0.25 / np.sqrt(8)
    return p.astype(np.float32)


def test_forward_matches_oracle(sw):
    x = np.random.default_rng(1).standard_normal((3, 1024)).astype(np.float32)
    p = _params(2)
    print("surrogate forward B=3", flush=True)
    with sw.Surrogate(1024, 4, seed=0) as nn:
        nn.set_params(p)
        assert np.array_equal(nn.get_params(), p)
        y = nn.forward(x)
    want, _ = no.forward(p.astype(np.float64), x, bf=True)
    assert _rel(y, want) < 3e-3, _rel(y, want)


def test_gradient_matches_oracle(sw):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 1024)).astype(np.float32)
    t = (0.3 * rng.standard_normal((2, 1024))).astype(np.float32)
    p = _params(4)
    print("surrogate gradient B=2", flush=True)
    with sw.Surrogate(1024, 2, seed=0) as nn:
        nn.set_params(p)
        loss, g = nn.gradient(x, t)
    wl, wg = no.loss_grad(p.astype(np.float64), x, t, bf=True)
    assert abs(loss - wl) < 1e-3 * wl
    # per parameter block: BF16 gradients between layers, FP32 accumulation
    o = 0
    for name, shape in no.SIZES:
        k = int(np.prod(shape))
        assert _rel(g[o:o + k], wg[o:o + k]) < 3e-2, (name, _rel(g[o:o + k], wg[o:o + k]))
        o += k


def test_training_reduces_error_and_judge(sw, wl):
    """Train on KS snapshots from the batched solver; the network's K-step
    change beats the 'no network' prediction du = 0 (relative error 1) on
    held-out trajectories, and the judged integrator falls back exactly when
    its tolerance demands."""
    B, nx, dx, dt, K = 32, 1024, 0.5, 0.005, 20
    eq = np.ones(B, np.int32)
    nxs = np.full(B, nx, np.int64)
    dts = np.full(B, dt)
    dxs = np.full(B, dx)
    ics = [wl.fourier_ic(nx, nx * dx, 100 + i) for i in range(B)]
    print("surrogate training", flush=True)
    with sw.Surrogate(nx, 64, seed=1) as nn:
        with sw.Batch(eq, nxs, dts, dxs) as b:
            b.set_state(np.concatenate(ics))
            b.advance(2000)  # past the smooth start
            nsamp, scales = nn.dataset(b, 60, K)
        assert nsamp == 60 * B and scales[0] > 0 and scales[1] > 0
        e0 = nn.eval(nsamp - 8 * B, 8 * B)
        loss = nn.fit(300, batch=32, lr=2e-3, nfit=nsamp - 8 * B)
        e1 = nn.eval(nsamp - 8 * B, 8 * B)
        print(f"held-out relative error of du: {e0:.3f} -> {e1:.3f} (loss {loss:.4f})", flush=True)
        assert e0 > 0.95 and e1 < 0.5
        # judged integration: tol = 0 falls back at every judged step, and with
        # judge_every = 1 the result is the plain solver's, bitwise
        with sw.Batch(eq[:4], nxs[:4], dts[:4], dxs[:4]) as b1, sw.Batch(eq[:4], nxs[:4], dts[:4], dxs[:4]) as b2:
            u0 = np.concatenate(ics[:4])
            b1.set_state(u0)
            b2.set_state(u0)
            st = nn.advance(b1, 3, K, judge_every=1, tol=0.0)
            b2.advance(3 * K)
            assert st["judged"] == 3 and st["fallbacks"] == 12
            assert np.array_equal(b1.get_state(), b2.get_state())
            st = nn.advance(b1, 5, K, judge_every=10, tol=1e-2)
            assert st["judged"] == 0 and st["network_steps"] == 5
