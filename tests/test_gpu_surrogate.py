"""The reflectance surrogate end to end on the device, against the method without it.

An ensemble of four 12-256-256-256-33 networks is trained (device kernels) on batched Monte Carlo
data of the two-layer-plus-vessel skin model (8192 samples x 20000 packets), then evaluated on 48
held-out samples re-simulated at 1e6 packets.  Plain Monte Carlo reaching the surrogate's median
error needs N* packets per query, with N* from the reference run's own standard errors (relative
SE ~ 1/sqrt(N)).  Bars, written here: median relative error <= 5% at this small training set
(the full-size pipeline reaches < 1% with 2e5 samples); surrogate queries/s at least 1000x plain
Monte Carlo queries/s at that matched error; the ensemble spread ranks the errors (the quarter of
queries with the largest spread has a larger median error than the rest); out-of-range queries
are always re-run by Monte Carlo and get its answer.
"""
import time

import numpy as np
import pytest

from ymc3d_amd import configs, surrogate as S, surrogate_data as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(engine):
    P, mc = D.generate(engine, 8192, 20_000, seed=11, progress_every=1e9)
    rate = mc["packets"] / mc["wall_s"]
    ens = S.Ensemble.train(P, S.encode_targets(mc["radial"], mc["total"]), members=4, epochs=40, batch=256, seed=5)
    yield ens, rate
    ens.close()


def test_surrogate_beats_plain_monte_carlo_at_matched_error(engine, trained):
    ens, mc_rate = trained
    Pt, ref = D.generate(engine, 48, 1_000_000, seed=12, tables_per_launch=8, progress_every=1e9)
    mean, spread = ens.predict(Pt)
    mc = np.concatenate([ref["radial"], ref["total"][:, None]], axis=1)
    se = np.concatenate([ref["radial_se"], ref["total_se"][:, None]], axis=1) / mc
    err = np.abs(10.0 ** mean - mc) / mc
    med = float(np.median(err))
    n_star = (np.median(se) * np.sqrt(1e6) / med) ** 2
    mc_queries_per_s = mc_rate / n_star
    Q = configs.sample_skin2_params(65536, 13)
    ens.predict(Q[:4096])
    t0 = time.perf_counter()
    ens.predict(Q)
    sur_queries_per_s = len(Q) / (time.perf_counter() - t0)
    print(f"median rel err {med:.4f} (MC ref rel SE {np.median(se):.4f}); plain MC needs {n_star:.3g} packets/query "
          f"-> {mc_queries_per_s:.3g} queries/s at {mc_rate:.3g} packets/s; surrogate {sur_queries_per_s:.3g} queries/s")
    assert med <= 0.05
    assert sur_queries_per_s >= 1000 * mc_queries_per_s
    big = spread.max(1) >= np.quantile(spread.max(1), 0.75)
    assert np.median(err[big]) > np.median(err[~big])


def test_out_of_range_queries_fall_back_to_monte_carlo(engine, trained):
    ens, _ = trained
    P = configs.sample_skin2_params(16, 21)
    P[:4, 0] = configs.PARAM_HI[0] * 3.0  # epidermal absorption far above the training range
    out = S.serve(ens, P, engine, threshold=1e9, fallback_photons=20_000, seed=3, tables_per_launch=16,
                  progress_every=1e9)
    assert out["flagged"][:4].all() and not out["flagged"][4:].any()
    assert np.array_equal(out["rows"], np.arange(4))
    mc = out["mc"]
    want = S.encode_targets(mc["radial"], mc["total"])
    np.testing.assert_array_equal(out["pred"][:4], want)
    np.testing.assert_array_equal(out["pred"][4:], out["mean"][4:])
