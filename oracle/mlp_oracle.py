"""CPU restatement of the warm-start network: forward, loss and backward.

TEST INFRASTRUCTURE ONLY.  tests/ use it as the checker of the GPU kernels;
the product never imports it.  The network has no counterpart in the
reference (whose start is MBPT2, setup_t_amplitudes, ccm_iter.f90:2329-2405):
it predicts the correction of the CCD amplitude over its MBPT2 value.

Network: 16 inputs -> 64 -> 64 -> 64 (GELU, the CDF as a polynomial) -> 1.
Parameters are one flat float vector in this order (row-major, W[k][o] maps
input k to output o):
    W1[16][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4[1]
Loss: mean over samples of (y - target)^2.

With bf16=True the functions round at the same points as the GPU kernels
(round-to-nearest-even to bfloat16): the weights of the three hidden layers,
the inputs, every hidden activation after GELU, and the pre-activation
gradients dz that feed the products.  Sums are float64 here and float32 on
the GPU.  Pinned by finite differences in tests/test_mlp.py.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 16, 64
SHAPES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NH)), ("b3", (NH,)), ("w4", (NH,)), ("b4", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round float values to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# GELU(z) = z Phi(z) with the normal CDF as the kernels define it: an odd
# polynomial on [-4, 4], Phi(z) = 1/2 + zc P(zc^2), zc = z clamped to [-4, 4]
# (coefficients of P in FP32, lowest order first).  tests/test_mlp.py pins it
# against the exact CDF (|error| <= 6e-5) and gelu_grad against differences.
PHI_C = np.array([3.988475204e-01, -6.617541611e-02, 9.664904326e-03, -1.048214734e-03,
                  8.066917508e-05, -4.101028480e-06, 1.217185570e-07, -1.580922393e-09],
                 np.float32).astype(np.float64)


def phi(z):
    zc = np.clip(z, -4.0, 4.0)
    return 0.5 + zc * np.polynomial.polynomial.polyval(zc * zc, PHI_C)


def gelu(z):
    return z * phi(z)


def gelu_grad(z):
    zc = np.clip(z, -4.0, 4.0)
    s = zc * zc
    p = np.polynomial.polynomial.polyval(s, PHI_C)
    dp = np.polynomial.polynomial.polyval(s, PHI_C[1:] * np.arange(1, len(PHI_C)))
    dphi = np.where(np.abs(z) < 4.0, p + 2.0 * s * dp, 0.0)
    return phi(z) + z * dphi


def forward(theta, x, bf=False):
    """x: (n, 16).  Returns y (n,) and the cache for backward."""
    p = unpack(theta)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    W1, W2, W3 = r(p["W1"]), r(p["W2"]), r(p["W3"])
    x = r(x)
    z1 = x @ W1 + p["b1"]
    h1 = r(gelu(z1))
    z2 = h1 @ W2 + p["b2"]
    h2 = r(gelu(z2))
    z3 = h2 @ W3 + p["b3"]
    h3 = r(gelu(z3))
    y = h3 @ p["w4"] + p["b4"][0]
    return y, (x, z1, h1, z2, h2, z3, h3, W2, W3, p)


def loss(theta, x, t, bf=False):
    y, _ = forward(theta, x, bf)
    return float(np.mean((y - t) ** 2))


def loss_grad(theta, x, t, bf=False):
    n = len(t)
    y, (x, z1, h1, z2, h2, z3, h3, W2, W3, p) = forward(theta, x, bf)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    g = 2.0 * (y - t) / n
    gw4 = h3.T @ g
    gb4 = np.array([np.sum(g)])
    dz3 = r(np.outer(g, p["w4"]) * gelu_grad(z3))
    gW3 = h2.T @ dz3
    gb3 = dz3.sum(0)
    dz2 = r((dz3 @ W3.T) * gelu_grad(z2))
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = r((dz2 @ W2.T) * gelu_grad(z1))
    gW1 = x.T @ dz1
    gb1 = dz1.sum(0)
    grads = pack({"W1": gW1, "b1": gb1, "W2": gW2, "b2": gb2, "W3": gW3, "b3": gb3,
                  "w4": gw4, "b4": gb4})
    return float(np.mean((y - t) ** 2)), grads
