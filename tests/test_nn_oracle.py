"""Pins the surrogate network's CPU oracle (oracle/nn_oracle.py): its hand-written
backward pass against central finite differences of its own float64 forward,
the periodic convolution against a direct loop, and the BF16 rounding rule
against known values.  CPU only."""
import numpy as np
import pytest

import nn_oracle as no


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(no.NPARAM) * scale / np.sqrt(64)


def test_param_count():
    assert no.NPARAM == 64 * 5 + 64 + 4 * (5 * 64 * 64 + 64) + 5 * 64 + 1 == 82881


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159265, 65504.0, 1e-40])
    got = no.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1+2^-7: even mantissa -> 1.0
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[2] == 1.0078125 and got[3] == 1.0078125
    assert got[4] == -3.140625
    assert got[5] == 65536.0


def test_forward_matches_direct_loops():
    p = _params(1)
    q = no.unpack(p)
    x = np.random.default_rng(2).standard_normal((2, 64))
    y, cache = no.forward(p, x, bf=False)
    X = 64
    # layer 1 and the output layer by explicit loops
    b, xi, c = 1, 5, 7
    z1 = q["b1"][c] + sum(q["w1"][c, t] * x[b, (xi + t - 2) % X] for t in range(5))
    assert abs(cache["Z"][0][b, xi, c] - z1) < 1e-12
    z3 = q["b3"][c] + sum(cache["A"][1][b, (xi + t - 2) % X, ci] * q["W3"][t, c, ci]
                          for t in range(5) for ci in range(64))
    assert abs(cache["Z"][2][b, xi, c] - z3) < 1e-12
    yo = q["b6"][0] + sum(q["w6"][t, ci] * cache["A"][4][b, (xi + t - 2) % X, ci]
                          for t in range(5) for ci in range(64))
    assert abs(y[b, xi] - yo) < 1e-12


def test_backward_matches_finite_differences():
    p = _params(3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 64))
    t = rng.standard_normal((2, 64)) * 0.1
    _, g = no.loss_grad(p, x, t, bf=False)
    idx = np.concatenate([rng.choice(np.arange(0, 384), 6, replace=False),          # L1
                          rng.choice(np.arange(384, 384 + 4 * 20544), 14, replace=False),
                          rng.choice(np.arange(no.NPARAM - 321, no.NPARAM), 5, replace=False)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (no.loss_grad(pp, x, t, bf=False)[0] - no.loss_grad(pm, x, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_gelu_tanh_form():
    """The oracle's GELU: z sigmoid(c (z + a z^3)) with c = 2 sqrt(2/pi),
    a = 0.044715.  Pinned against its limits, the erf form it approximates
    (within 5e-4, which a wrong constant exceeds) and, for the derivative,
    central differences."""
    from scipy.special import erf

    z = np.linspace(-8.0, 8.0, 4001)
    g = no.gelu(z)
    assert no.gelu(np.array([0.0]))[0] == 0.0
    assert np.allclose(no.gelu(np.array([30.0, -30.0])), [30.0, 0.0], atol=1e-12)
    exact = 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))
    assert np.max(np.abs(g - exact)) < 5e-4
    assert np.max(np.abs(g - exact)) > 1e-5  # it is the tanh form, not erf
    h = 1e-5
    fd = (no.gelu(z + h) - no.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - no.gelu_d(z))) < 1e-8
