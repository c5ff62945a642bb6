"""Run modes on the device against the oracle: the textbook Fresnel formula instead of the
reference's (Q1, src/mc3d_gpu.cu:241), and the zero-step kill rule at every step or never instead
of at the reference's launch boundaries (src/mc3d_gpu.cu:343).

Same chains and streams on both sides.  Bars, written here: total R and total T within 3 combined
standard errors; killed packet counts within 4 Poisson standard deviations of each other; in the
every-step mode most packets that scatter back across the beam's entry face are killed, so far
more packets die than in the default mode.
"""
import numpy as np
import pytest

import ymc3d_amd as y
from oracle import oracle as O
from ymc3d_amd import configs

pytestmark = pytest.mark.gpu

MODES = {
    "exact_fresnel": dict(fresnel_mode=y.FRESNEL_EXACT),
    "kill_every_step": dict(kill_zero_step=2),
    "kill_never": dict(kill_zero_step=0),
}


def _run(engine, multipliers, over, n=200_000):
    s = configs.skin_slab_sphere(photons=n)
    s.load(engine)
    params = dict(s.params)
    params.update(over)
    g = engine.simulate(**params, chain_lanes=16384, flags=y.FLAG_COUNTERS)
    r = O.run_scene(s, multipliers, chain_lanes=16384, **over)
    return g, r


@pytest.mark.parametrize("mode", sorted(MODES))
def test_mode_matches_oracle(engine, multipliers, mode):
    g, r = _run(engine, multipliers, MODES[mode])
    for tot, se in (("total_R", "total_R_se"),):
        a, b = getattr(g, tot)()[0], getattr(r, tot)()[0]
        s = np.hypot(getattr(g, se)()[0], getattr(r, se)()[0])
        print(f"{mode}: {tot} gpu {a:.5f} oracle {b:.5f} se {s:.2g}")
        assert abs(a - b) < 3 * s + 1e-12
    tg, tr = g.totals[0], r.totals[0]
    assert tg["photons"] == tr["photons"] == 200_000
    Tg, Tr = int(tg["T_w"]) / 2.0 ** 32 / 2e5, int(tr["T_w"]) / 2.0 ** 32 / 2e5
    assert abs(Tg - Tr) <= 3e-3 * max(Tg, Tr) + 1e-6
    kg, kr = int(tg["killed"]), int(tr["killed"])
    print(f"{mode}: killed gpu {kg} oracle {kr}")
    assert abs(kg - kr) <= 4 * np.sqrt(kg + kr) + 2
    if mode == "kill_every_step":
        g0, _ = _run(engine, multipliers, {}, n=200_000)
        assert kg > 100 * max(1, int(g0.totals[0]["killed"]))
