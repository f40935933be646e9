"""CPU oracle of the learned KS surrogate network (numpy).

TEST INFRASTRUCTURE: the checker for the GPU surrogate kernels
(1dsweptcuda_amd/csrc/sw1d_nn.hip); nothing in the library uses it.

Same network and parameter layout as the device code:
  L1   Z1[b,x,c] = b1[c] + sum_t w1[c,t] x[b,x+t-2]             (w1: [64][5])
  L2-5 Z[b,x,co] = b[co] + sum_t sum_ci A[b,x+t-2,ci] W[t,co,ci] (W: [5][64][64])
  L6   y[b,x]    = b6 + sum_t sum_c w6[t,c] A5[b,x+t-2,c]        (w6: [5][64])
  A = gelu(Z) (tanh form), indices periodic in x.
The backward pass is written out by hand (the reverse of each sum above).

With bf16=True the oracle applies the device's precision choices: BF16
weights of L2-5, BF16 activations, BF16 gradients dZ between layers, FP32
elsewhere (computed here in float64).  With bf16=False it is the plain float64
network, which the finite-difference check pins.
"""
from __future__ import annotations

import numpy as np

C, TAPS = 64, 5
SIZES = [("w1", (C, TAPS)), ("b1", (C,))]
for _l in range(2, 6):
    SIZES += [(f"W{_l}", (TAPS, C, C)), (f"b{_l}", (C,))]
SIZES += [("w6", (TAPS, C)), ("b6", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


GELU_C = 2.0 * np.sqrt(2.0 / np.pi)
GELU_A = 0.044715


def _gelu_sigmoid(z):
    with np.errstate(over="ignore"):
        return 1.0 / (1.0 + np.exp(-GELU_C * (z + GELU_A * z ** 3)))


def gelu(z):
    """GELU in its tanh form: z sigmoid(c (z + a z^3)), c = 2 sqrt(2/pi),
    a = 0.044715 (Hendrycks & Gimpel), the form the device evaluates."""
    return z * _gelu_sigmoid(z)


def gelu_d(z):
    s = _gelu_sigmoid(z)
    return s + z * s * (1.0 - s) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def _sh(a, t):
    """a[b, x + t - 2] along axis 1 (periodic)."""
    return np.roll(a, -(t - 2), axis=1)


def forward(p, x, bf=True):
    q = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    x = np.asarray(x, np.float64)
    Z = [q["b1"] + sum(_sh(x, t)[..., None] * q["w1"][:, t] for t in range(TAPS))]
    A = [r(gelu(Z[0]))]
    for l in range(2, 6):
        W = r(q[f"W{l}"])
        z = q[f"b{l}"] + sum(_sh(A[-1], t) @ W[t].T for t in range(TAPS))
        Z.append(z)
        A.append(r(gelu(z)))
    y = q["b6"][0] + sum(_sh(A[-1], t) @ q["w6"][t] for t in range(TAPS))
    return y, {"x": x, "Z": Z, "A": A}


def loss_grad(p, x, target, bf=True):
    q = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    y, cache = forward(p, x, bf)
    N = y.size
    d = y - np.asarray(target, np.float64)
    loss = float(np.mean(d * d))
    g = 2.0 * d / N
    A, Z, x = cache["A"], cache["Z"], cache["x"]
    grad = {}
    grad["w6"] = np.stack([np.einsum("bx,bxc->c", g, _sh(A[4], t)) for t in range(TAPS)])
    grad["b6"] = np.array([g.sum()])
    dA = sum(np.roll(g, t - 2, axis=1)[..., None] * q["w6"][t] for t in range(TAPS))
    dZ = r(dA * gelu_d(Z[4]))
    for l in range(5, 1, -1):
        W = r(q[f"W{l}"])
        Ain = A[l - 2]
        grad[f"W{l}"] = np.stack([np.einsum("bxo,bxi->oi", dZ, _sh(Ain, t)) for t in range(TAPS)])
        grad[f"b{l}"] = dZ.sum(axis=(0, 1))
        dA = sum(np.roll(dZ, t - 2, axis=1) @ W[t] for t in range(TAPS))
        dZ = r(dA * gelu_d(Z[l - 2]))
    grad["w1"] = np.stack([np.einsum("bxc,bx->c", dZ, _sh(x, t)) for t in range(TAPS)], axis=1)
    grad["b1"] = dZ.sum(axis=(0, 1))
    return loss, pack(grad)
