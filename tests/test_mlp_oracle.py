"""The CPU MLP reference (oracle/mlp_oracle.py) pinned by central finite differences in float64,
and its GELU by the closed form z*Phi(z) at known points."""
import numpy as np
import pytest

from oracle import mlp_oracle as R


def test_gelu_values():
    # Phi(1) = 0.841344746..., Phi(-1) = 0.158655254...
    assert R.gelu(np.array([0.0, 1.0, -1.0])) == pytest.approx([0.0, 0.8413447460685429, -0.15865525393145707])
    z = np.linspace(-4, 4, 41)
    h = 1e-6
    assert R.gelu_grad(z) == pytest.approx((R.gelu(z + h) - R.gelu(z - h)) / (2 * h), abs=1e-7)


def test_gradient_finite_differences():
    rng = np.random.default_rng(0)
    widths = (5, 7, 6, 3)
    n = sum(a * b + b for a, b in zip(widths[:-1], widths[1:]))
    p = rng.normal(size=n) * 0.5
    x = rng.normal(size=(9, 5))
    y = rng.normal(size=(9, 3))
    _, g = R.loss_and_grad(p, widths, x, y)
    h = 1e-6
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd = (R.loss_and_grad(p + e, widths, x, y)[0] - R.loss_and_grad(p - e, widths, x, y)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def test_device_gelu_formula():
    """The device's GELU (mlp.hip: gelu, gelu_grad), restated in float64: z Phi(z) = max(z, 0) - q |z|
    with q = erfc(|z|/sqrt 2)/2 from Abramowitz-Stegun 7.1.26, in the scaled variable
    b = |z| sqrt(log2(e)/2).  Against the exact erf: GELU within 2.2e-7 (and 7e-8 |z|), its derivative
    within 3e-7, over [-12, 12]; the constants are the published ones rescaled."""
    from math import log
    from scipy.special import erf

    log2e = 1 / log(2)
    k1, pp = np.sqrt(log2e / 2), 0.3275911 / np.sqrt(log2e)
    assert k1 == pytest.approx(0.84932180028801907, rel=1e-15) and pp == pytest.approx(0.27273748087922250, rel=1e-15)
    c = [0.5 * v for v in (0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429)]
    assert c == pytest.approx([0.127414796, -0.142248368, 0.7107068705, -0.7265760135, 0.5307027145], rel=1e-12)
    z = np.linspace(-12, 12, 400001)
    b = np.abs(z) * k1
    t = 1 / (1 + pp * b)
    e = np.exp2(-b * b)
    q = t * (c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * c[4])))) * e
    g = np.maximum(z, 0) - q * np.abs(z)
    d = np.where(z >= 0, 1 - q, q) + z * 0.39894228040143268 * e
    assert np.max(np.abs(g - R.gelu(z))) < 2.2e-7
    assert np.all(np.abs(g - R.gelu(z)) <= 7e-8 * np.abs(z) + 1e-15)
    assert np.max(np.abs(d - R.gelu_grad(z))) < 3e-7
