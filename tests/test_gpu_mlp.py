"""Surrogate MLP kernels on the GPU against the CPU FP32 reference (oracle/mlp_oracle.py).

The device multiplies BF16 operands (weights, activations and back-propagated errors rounded to
bfloat16, 8 significant bits) with FP32 accumulation.  Bars, written here: forward outputs within
3% relative L2 of the FP32 network; every parameter tensor's gradient with cosine similarity >=
0.995 and relative L2 error <= 6% against the FP32 gradient; odd sizes (12 inputs, 33 outputs,
batches not a multiple of the 64-row tile) exercise the edge handling of all three GEMM layouts.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as R
from ymc3d_amd import surrogate as S

pytestmark = pytest.mark.gpu


def _bf16(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.fixture(scope="module")
def net():
    n = S.MLP(seed=7)
    yield n
    n.close()


def test_forward_matches_fp32(net):
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (300, 12)).astype(np.float32)
    p = net.get_params()
    got = net.forward(x)
    want, _, _ = R.forward(_bf16(p).astype(np.float64), S.WIDTHS, _bf16(x))
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert got.shape == (300, 33) and rel < 0.03, rel


def test_gradients_match_fp32(net):
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (200, 12)).astype(np.float32)
    y = rng.normal(size=(200, 33)).astype(np.float32)
    p = net.get_params()
    loss, g = net.gradients(x, y)
    rloss, rg = R.loss_and_grad(_bf16(p).astype(np.float64), S.WIDTHS, _bf16(x), y)
    assert loss == pytest.approx(rloss, rel=0.03)
    for (gW, gb), (rW, rb) in zip(R.unpack(g.astype(np.float64), S.WIDTHS), R.unpack(rg, S.WIDTHS)):
        for a, b in ((gW, rW), (gb, rb)):
            cos = float(a.ravel() @ b.ravel() / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
            rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
            assert cos >= 0.995 and rel <= 0.06, (a.shape, cos, rel)


def test_training_reduces_loss_and_fits_a_smooth_map():
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (8192, 12)).astype(np.float32)
    A = rng.normal(size=(12, 33)) / 4
    y = (np.tanh(x @ A) + 0.1 * np.sin(3 * x[:, :1])).astype(np.float32)
    net = S.MLP(seed=3)
    before = float(np.mean((net.forward(x) - y) ** 2))
    losses = net.fit(x, y, epochs=30, batch=256, lr=2e-3)
    pred = net.forward(x)
    net.close()
    after = float(np.mean((pred - y) ** 2))
    print(f"training MSE {before:.4f} -> {after:.6f}; last-batch losses {losses[0]:.5f} -> {losses[-1]:.6f}")
    assert after < 0.02 * before and after < 2e-3
    assert losses[-1] < losses[0]


def test_fused_ensemble_forward():
    """The fused ensemble pass (every layer in one launch, 32x32x16 BF16 products, rows chunked by
    4096) against the FP32 reference for each member, against each member alone, and against the
    layer-by-layer products of the training path (same loss on a batch within 0.2%)."""
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (5000, 12)).astype(np.float32)
    y = rng.normal(size=(5000, 33)).astype(np.float32)
    nets = [S.MLP(seed=s) for s in (11, 12, 13)]
    try:
        ens = S.forward_ensemble(nets, x)
        assert ens.shape == (3, 5000, 33)
        for q, n in enumerate(nets):
            np.testing.assert_array_equal(ens[q], n.forward(x))
            want, _, _ = R.forward(_bf16(n.get_params()).astype(np.float64), S.WIDTHS, _bf16(x))
            rel = np.linalg.norm(ens[q] - want) / np.linalg.norm(want)
            assert rel < 0.03, rel
            loss, _ = n.gradients(x[:512], y[:512])
            assert float(np.mean((ens[q][:512] - y[:512]) ** 2)) == pytest.approx(loss, rel=2e-3)
    finally:
        for n in nets:
            n.close()


def test_fused_single_layer_rounds_operands_to_nearest_even():
    """One linear layer (12 -> 33): the fused pass computes bf16(x) bf16(W) + b with FP32 accumulation, so
    its outputs match float64 products of the RNE-rounded operands to FP32 accuracy; a truncating or
    otherwise different conversion would be off by ~1e-3."""
    rng = np.random.default_rng(8)
    net = S.MLP((12, 33), seed=5)
    try:
        p = rng.normal(size=net.n_params).astype(np.float32)
        net.set_params(p)
        x = rng.uniform(-3, 3, (1000, 12)).astype(np.float32)
        got = net.forward(x)
        W, b = p[:12 * 33].reshape(12, 33), p[12 * 33:]
        xb, Wb = _bf16(x).astype(np.float64), _bf16(W).astype(np.float64)
        want = xb @ Wb + b
        scale = np.abs(xb) @ np.abs(Wb) + np.abs(b)  # FP32 accumulation error is relative to this
        assert np.max(np.abs(got - want) / scale) < 2e-6
        # operands truncated instead of rounded would leave errors of order 1e-3 of the scale
        xt = (x.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32).astype(np.float64)
        assert np.max(np.abs(xt @ Wb + b - want) / scale) > 1e-4
    finally:
        net.close()


def test_fused_gelu_equals_exact_gelu_in_bf16():
    """The device GELU through the fused pass: layer 1 copies input 0 to 32 hidden units (weights 1,
    bias 0), layer 2 copies hidden unit i to output i (weights 1), so output i = bf16(GELU(bf16(x0))).
    Against the exact erf-based GELU rounded to BF16: within one BF16 unit or 1e-6 everywhere (the
    device's erfc approximation has absolute error <= 1.5e-7, so far in the negative tail, where GELU
    is below 1e-6, only the absolute error is small), and equal for at least 99.5% of the inputs
    with |GELU| >= 1e-3."""
    from scipy.special import erf

    net = S.MLP((12, 32, 33), seed=1)
    try:
        W1 = np.zeros((12, 32), np.float32)
        W1[0, :] = 1.0
        W2 = np.zeros((32, 33), np.float32)
        W2[np.arange(32), np.arange(32)] = 1.0
        net.set_params(np.concatenate([W1.ravel(), np.zeros(32, np.float32), W2.ravel(), np.zeros(33, np.float32)]))
        z = _bf16(np.linspace(-9, 9, 20001).astype(np.float32))
        x = np.zeros((z.size, 12), np.float32)
        x[:, 0] = z
        got = net.forward(x)[:, :32]
        assert np.all(got == got[:, :1])
        zd = z.astype(np.float64)
        want = _bf16((0.5 * zd * (1 + erf(zd / np.sqrt(2)))).astype(np.float32))
        g = got[:, 0]
        gu, wu = (v.view(np.int32).astype(np.int64) >> 16 for v in (g, want))  # BF16 bit patterns
        close = ((np.abs(gu - wu) <= 1) & (np.sign(g) == np.sign(want))) | (np.abs(g - want) <= 1e-6)
        assert np.all(close), z[~close][:10]
        big = np.abs(want) >= 1e-3
        assert np.mean(g[big] == want[big]) >= 0.995, np.mean(g[big] == want[big])
    finally:
        net.close()


def test_concurrent_member_training_equals_sequential():
    """Ensemble members train from concurrent host threads, each network on its own stream: every
    member's parameters equal those of the same member trained alone."""
    rng = np.random.default_rng(7)
    P = np.column_stack([rng.uniform(lo, hi, 3000) for lo, hi in zip(S.configs.PARAM_LO, S.configs.PARAM_HI)])
    targets = rng.normal(size=(3000, S.WIDTHS[-1])).astype(np.float32)
    ens = S.Ensemble.train(P, targets, members=3, epochs=2, batch=256, seed=5)
    try:
        x = S.encode_params(P)
        yz = ((targets - ens.y_mean) / ens.y_std).astype(np.float32)
        for m, member in enumerate(ens.members):
            alone = S.MLP(seed=5 * 1000 + m)
            idx = np.random.default_rng(5 * 1000 + m).integers(0, len(x), len(x))
            alone.fit(x[idx], yz[idx], 2, 256, 2e-3, 2e-5, 5 * 1000 + m)
            assert np.array_equal(alone.get_params(), member.get_params()), m
            alone.close()
    finally:
        ens.close()
