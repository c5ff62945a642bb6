"""GPU: ymc_mlp_forward_ensemble_noalloc, the fused ensemble pass that allocates nothing, against
ymc_mlp_forward_ensemble through the C ABI.  Both run the same kernel on the same packed weights,
so the outputs must be identical bit for bit.  The no-allocation pass runs first on fresh networks
(it must not depend on buffers an earlier allocating call left behind), then twice more, the last
call smaller than the one before, so that stale rows of the shared staging buffer would show."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(lib, nets, x):
    from ymc3d_amd import surrogate as S
    from ymc3d_amd._abi import ptr
    from ymc3d_amd.engine import check

    y = np.full((len(nets), x.shape[0], nets[0].widths[-1]), np.nan, np.float32)
    hp, keep = S._handles(nets)
    check(lib.ymc_mlp_forward_ensemble_noalloc(hp, len(nets), ptr(x, C.c_float), x.shape[0], ptr(y, C.c_float)),
          "ymc_mlp_forward_ensemble_noalloc")
    return y


def test_noalloc_ensemble_pass_equals_the_allocating_pass():
    from ymc3d_amd import surrogate as S
    from ymc3d_amd.engine import load_library

    lib = load_library()
    nets = [S.MLP(S.WIDTHS, seed=11 + s) for s in range(4)]
    try:
        rng = np.random.default_rng(5)
        for rows, change_params in ((9000, False), (5000, True), (333, False)):
            print(f"noalloc pass: {rows} rows, params changed {change_params}", flush=True)
            if change_params:  # the packed weights are rebuilt in place
                p = nets[2].get_params()
                nets[2].set_params(p * np.float32(0.5))
            x = rng.standard_normal((rows, S.WIDTHS[0])).astype(np.float32)
            got = _noalloc(lib, nets, x)
            want = S.forward_ensemble(nets, x)
            assert np.all(np.isfinite(got))
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    finally:
        for n in nets:
            n.close()


def test_noalloc_single_network_and_empty_input():
    from ymc3d_amd import surrogate as S
    from ymc3d_amd.engine import load_library

    lib = load_library()
    net = S.MLP((12, 64, 33), seed=3)
    try:
        x = np.random.default_rng(6).standard_normal((100, 12)).astype(np.float32)
        got = _noalloc(lib, [net], x)
        assert np.array_equal(got[0].view(np.uint32), net.forward(x).view(np.uint32))
        assert _noalloc(lib, [net], x[:0]).shape == (1, 0, 33)
    finally:
        net.close()
