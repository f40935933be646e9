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


def test_forward_matches_oracle_unfused_width(sw):
    """nx = 192 is not a multiple of the 256-position tile: the per-layer
    64-position kernels run instead of the fused ones."""
    x = np.random.default_rng(5).standard_normal((2, 192)).astype(np.float32)
    p = _params(6)
    print("surrogate forward nx=192 B=2", flush=True)
    with sw.Surrogate(192, 2, seed=0) as nn:
        nn.set_params(p)
        y = nn.forward(x)
    want, _ = no.forward(p.astype(np.float64), x, bf=True)
    assert _rel(y, want) < 3e-3, _rel(y, want)


def test_network_step_is_forward_plus_update(sw):
    """One unjudged network step of the integrator (fused kernels reading the
    FP64 state) equals u + out_scale * net(FP32(u) * FP32(1/in_scale)) from
    the forward entry point, bit for bit."""
    import importlib

    wl = importlib.import_module("1dsweptcuda_amd.workloads")
    B, nx = 3, 1024
    u0 = np.concatenate([wl.fourier_ic(nx, nx * 0.5, 40 + i) for i in range(B)])
    p = _params(7)
    in_s, out_s = 1.7, 0.05
    print("surrogate network step B=3", flush=True)
    with sw.Surrogate(nx, B, seed=0) as nn, sw.Batch(np.ones(B, np.int32), np.full(B, nx, np.int64),
                                                      np.full(B, 0.005), np.full(B, 0.5)) as b:
        nn.set_params(p)
        nn.set_scales(in_s, out_s)
        b.set_state(u0)
        st = nn.advance(b, 1, 20, judge_every=0)
        assert st["network_steps"] == 1 and st["judged"] == 0
        u1 = b.get_state()
        x = (u0.astype(np.float32) * np.float32(1.0 / in_s)).reshape(B, nx)
        y = nn.forward(x)
    want = u0 + out_s * y.ravel().astype(np.float64)
    assert np.array_equal(u1, want), float(np.abs(u1 - want).max())


def test_whole_network_kernel_matches_per_layer_launches(sw):
    """At nx = 1024 inference is one whole-network launch (the activation
    resident in LDS); the five per-layer launches give the same bits, for the
    forward entry point and for a network step on FP64 states."""
    import importlib

    wl = importlib.import_module("1dsweptcuda_amd.workloads")
    B, nx = 5, 1024
    x = np.random.default_rng(11).standard_normal((B, nx)).astype(np.float32)
    u0 = np.concatenate([wl.fourier_ic(nx, nx * 0.5, 60 + i) for i in range(B)])
    p = _params(12)
    out = {}
    with sw.Surrogate(nx, B, seed=0) as nn:
        nn.set_params(p)
        nn.set_scales(1.3, 0.07)
        for mode in (nn.INFER_WHOLE, nn.INFER_PER_LAYER):
            print("surrogate inference mode", mode, flush=True)
            nn.set_inference(mode)
            y = nn.forward(x)
            with sw.Batch(np.ones(B, np.int32), np.full(B, nx, np.int64), np.full(B, 0.005),
                          np.full(B, 0.5)) as b:
                b.set_state(u0)
                nn.advance(b, 3, 20, judge_every=0)
                out[mode] = (y, b.get_state())
    assert np.array_equal(out[1][0], out[2][0])
    assert np.array_equal(out[1][1], out[2][1])
    assert not np.array_equal(out[2][1], u0)


def test_inference_mode_validation(sw):
    with sw.Surrogate(512, 2, seed=0) as nn:
        with pytest.raises(sw.SweptError):
            nn.set_inference(nn.INFER_WHOLE)  # the whole-network kernel is for nx = 1024
        nn.set_inference(nn.INFER_PER_LAYER)
    with sw.Surrogate(192, 2, seed=0) as nn:
        with pytest.raises(sw.SweptError):
            nn.set_inference(nn.INFER_PER_LAYER)  # not a multiple of 256


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


def test_dataset_pairs_and_scales_match_oracle(sw, orc, wl):
    """sw1d_nn_dataset: snapshots every K steps of an FP64 KS batch; in_scale
    and out_scale are the RMS of u and of the K-step change over the
    snapshots, and each stored pair is (u_s / in_scale, (u_{s+1} - u_s) /
    out_scale) rounded once to float.  Snapshots from the oracle's KS run."""
    B, nx, K, nsnap, lead = 4, 1024, 20, 5, 100
    ics = [wl.fourier_ic(nx, nx * 0.5, 300 + i) for i in range(B)]
    c = orc.ks_consts(0.005)
    snaps = []  # [nsnap + 1][B][nx]
    u = [orc.ks_run(u0, lead, c) for u0 in ics]
    snaps.append(np.array(u))
    for _ in range(nsnap):
        u = [orc.ks_run(v, K, c) for v in u]
        snaps.append(np.array(u))
    snaps = np.array(snaps)
    d = np.diff(snaps, axis=0)
    in_scale = np.sqrt(np.mean(snaps[:-1] ** 2))
    out_scale = np.sqrt(np.mean(d ** 2))
    print("surrogate dataset B=4 nsnap=5", flush=True)
    with sw.Surrogate(nx, 8, seed=0) as nn, sw.Batch(np.ones(B, np.int32), np.full(B, nx, np.int64),
                                                      np.full(B, 0.005), np.full(B, 0.5)) as b:
        b.set_state(np.concatenate(ics))
        b.advance(lead)
        n, scales = nn.dataset(b, nsnap, K)
        assert n == nsnap * B
        x, t = nn.dataset_samples(0, n)
    assert abs(scales[0] / in_scale - 1) < 1e-12 and abs(scales[1] / out_scale - 1) < 1e-12
    for s in range(nsnap):
        for p in range(B):
            k = s * B + p
            assert np.array_equal(x[k], (snaps[s, p] * (1.0 / scales[0])).astype(np.float32)), (s, p)
            assert np.array_equal(t[k], ((snaps[s + 1, p] - snaps[s, p]) * (1.0 / scales[1])).astype(np.float32))


def test_judge_keeps_some_and_replaces_others(sw, wl):
    """A tolerance between the smallest and largest per-problem error of one
    network step: each problem then holds either the network's state (error
    within tolerance) or the plain solver's (fallback), bit for bit, and the
    fallback count and largest judged error agree with the distances."""
    B, nx, K = 8, 1024, 20
    rng = np.random.default_rng(21)
    p = (rng.standard_normal(sw.Surrogate.param_count()) * 0.05).astype(np.float32)
    # uneven amplitudes give uneven relative errors
    u0 = np.concatenate([(0.3 + 0.4 * i) * wl.fourier_ic(nx, nx * 0.5, 500 + i) for i in range(B)])
    args = (np.ones(B, np.int32), np.full(B, nx, np.int64), np.full(B, 0.005), np.full(B, 0.5))
    print("surrogate mixed judge B=8", flush=True)
    with sw.Surrogate(nx, B, seed=0) as nn, sw.Batch(*args) as b, sw.Batch(*args) as ref:
        nn.set_params(p)
        nn.set_scales(1.0, 0.02)
        b.set_state(u0)
        nn.advance(b, 1, K, judge_every=0)
        net = b.get_state().reshape(B, nx)
        ref.set_state(u0)
        ref.advance(K)
        plain = ref.get_state().reshape(B, nx)
        e = np.linalg.norm(net - plain, axis=1) / np.linalg.norm(plain, axis=1)
        order = np.sort(e)
        assert order[0] < order[-1]
        gaps = np.diff(order)
        j = int(np.argmax(gaps[1:-1])) + 1 if B > 3 else int(np.argmax(gaps))
        tol = 0.5 * (order[j] + order[j + 1])  # far from every distance
        b.set_state(u0)
        st = nn.advance(b, 1, K, judge_every=1, tol=tol)
        got = b.get_state().reshape(B, nx)
    kept = e <= tol
    assert 0 < kept.sum() < B
    assert st["judged"] == 1 and st["fallbacks"] == int((~kept).sum())
    assert abs(st["max_judged_error"] / e.max() - 1) < 1e-9
    for i in range(B):
        assert np.array_equal(got[i], net[i] if kept[i] else plain[i]), (i, e[i], tol)
