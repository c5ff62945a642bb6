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
