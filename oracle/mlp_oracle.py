"""CPU FP32 reference of the reflectance-surrogate MLP (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the ymc3d_amd package.

Network: widths w0..wL; hidden layers z = h W + b, h' = GELU(z) with GELU(z) = z Phi(z) (exact,
erf-based); output layer linear; loss = mean over all outputs of (out - y)^2.  Parameters are one
flat vector, per layer W [in x out] row-major then b [out].  Gradients by reverse-mode
differentiation in float64, the reference for the device's BF16-operand / FP32-accumulate
kernels.  The GELU derivative is Phi(z) + z phi(z).
"""
from __future__ import annotations

import numpy as np
from math import sqrt

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(__import__("math").erf)


def unpack(params: np.ndarray, widths):
    out, o = [], 0
    for a, b in zip(widths[:-1], widths[1:]):
        W = params[o:o + a * b].reshape(a, b)
        o += a * b
        c = params[o:o + b]
        o += b
        out.append((W, c))
    assert o == params.size
    return out


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / sqrt(2.0))) + z * np.exp(-0.5 * z * z) / sqrt(2.0 * np.pi)


def forward(params, widths, x):
    layers = unpack(np.asarray(params, np.float64), widths)
    h = np.asarray(x, np.float64)
    zs, hs = [], [h]
    for l, (W, b) in enumerate(layers):
        z = h @ W + b
        zs.append(z)
        h = gelu(z) if l + 1 < len(layers) else z
        hs.append(h)
    return h, zs, hs


def loss_and_grad(params, widths, x, y):
    out, zs, hs = forward(params, widths, x)
    y = np.asarray(y, np.float64)
    n = out.size
    loss = float(np.mean((out - y) ** 2))
    layers = unpack(np.asarray(params, np.float64), widths)
    grads = []
    dz = 2.0 * (out - y) / n
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads.append((hs[l].T @ dz, dz.sum(0)))
        if l == 0:
            break
        dh = dz @ W.T
        dz = dh * gelu_grad(zs[l - 1])
    flat = []
    for gW, gb in reversed(grads):
        flat += [gW.ravel(), gb]
    return loss, np.concatenate(flat)
