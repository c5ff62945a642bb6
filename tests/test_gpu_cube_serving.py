"""Serving the full hyperspectral cube with Monte Carlo fallback (GPU).

Workload: 256 x 256 pixels x 64 wavelengths = 4 194 304 parameter queries of the two-layer-plus-vessel
skin model, 5% of the pixels drawn from widened ranges (surrogate_data.hyperspectral_cube), answered
by the shipped 4-member ensemble (surrogate.DEFAULT_ENSEMBLE) in fused passes of 4096 queries.
Queries outside the training ranges or with ensemble spread above the ensemble's stored threshold
are flagged for 1e6-packet Monte Carlo, and a seeded 1% of the others is audited by Monte Carlo.
The Monte Carlo runs on seeded samples of 128 flagged and 256 audit queries
(surrogate_pipeline.measure_cube); a 4 x 4-pixel cube runs through surrogate.serve in full.

Bars, written here: every out-of-range query flagged; audit share of the unflagged queries 1% within
4 binomial standard deviations, and no audit query among the flagged; the triage means equal an
independent ensemble pass bit for bit; served answers are the ensemble mean where unflagged and the
Monte Carlo answer where flagged; the audit's median relative error, less the audit Monte Carlo's
own noise in quadrature, at most 1%; serving with fallback faster than plain Monte Carlo at the
surrogate's accuracy.
"""
import numpy as np
import pytest

from ymc3d_amd import surrogate as S, surrogate_data as D, surrogate_pipeline as SP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(engine):
    ens, extra = S.Ensemble.load(S.DEFAULT_ENSEMBLE)
    print("256x256x64 cube, fallback sample 128, audit sample 256", flush=True)
    res = SP.measure_cube(ens, engine, float(extra["threshold"]), float(extra["n_star"]), cube=(256, 256),
                          fallback_sample=128, audit_sample=256, seed=0)
    keys = ("queries", "flagged_frac", "out_of_range", "audit", "surrogate_queries_per_s", "fallback_s_per_query",
            "spectra_per_s_with_fallback", "audit_median_rel_err", "audit_mc_median_rel_se",
            "audit_rel_err_less_mc_noise", "mc_spectra_per_s_at_matched_error", "speedup_with_fallback")
    print({k: res[k] for k in keys}, flush=True)
    yield ens, extra, res
    ens.close()


def test_every_out_of_range_query_is_flagged(served):
    _, _, r = served
    Q, flagged, outl = r["Q"], r["flagged_mask"], r["outlier_mask"]
    assert r["queries"] == 256 * 256 * 64 == len(Q)
    inr = S.in_range(Q)
    assert flagged[~inr].all()
    assert (~inr).sum() == r["out_of_range"] > 0
    # outlier pixels were drawn from widened ranges; in-range pixels were clipped into the ranges
    assert inr[~outl].all() and np.mean(~inr[outl]) > 0.5
    assert abs(outl.mean() - 0.05) < 4 * np.sqrt(0.05 * 0.95 / (256 * 256))


def test_audit_share_is_one_percent_of_the_unflagged(served):
    _, _, r = served
    flagged, audit = r["flagged_mask"], r["audit_mask"]
    assert not (audit & flagged).any()
    n = int((~flagged).sum())
    p = audit[~flagged].mean()
    assert abs(p - 0.01) <= 4 * np.sqrt(0.01 * 0.99 / n), (p, n)


def test_triage_means_equal_an_independent_ensemble_pass(served):
    ens, _, r = served
    idx = np.sort(np.random.default_rng(9).choice(r["queries"], 4096, replace=False))
    outs = S.forward_ensemble(ens.members, S.encode_params(r["Q"][idx])) * ens.y_std + ens.y_mean
    np.testing.assert_array_equal(outs.mean(0), r["mean"][idx])
    np.testing.assert_array_equal(outs.std(0), r["spread"][idx])


def test_served_answers_on_a_small_cube(engine, served):
    ens, extra, _ = served
    Q, _ = D.hyperspectral_cube(4, 4, 64, 0.25, seed=77)
    out = S.serve(ens, Q, engine, float(extra["threshold"]), 20_000, audit_frac=0.05, seed=3, tables_per_launch=512,
                  progress_every=1e9)
    flagged, rows = out["flagged"], out["rows"]
    assert flagged.any() and (~flagged).any()
    np.testing.assert_array_equal(out["pred"][~flagged], out["mean"][~flagged])
    mc = out["mc"]
    fb = flagged[rows]
    np.testing.assert_array_equal(out["pred"][rows[fb]], S.encode_targets(mc["radial"][fb], mc["total"][fb]))
    assert np.array_equal(rows, np.flatnonzero(flagged | out["audit"]))


def test_fallback_and_audit_monte_carlo(served):
    _, _, r = served
    fb, au = r["fallback_mc"], r["audit_mc"]
    assert r["fallback_sample"] == 128 and r["audit_sample"] == 256
    assert fb["packets"] == 128 * 1_000_000 and au["packets"] == 256 * 1_000_000
    for m in (fb, au):
        assert np.all(np.isfinite(m["total"])) and np.all(m["total"] > 0)
    # the surrogate's own error, once the 1e6-packet audit's noise is removed, meets the 1% target
    assert r["audit_mc_median_rel_se"] < 0.02
    assert r["audit_rel_err_less_mc_noise"] <= 0.01
    assert r["speedup_with_fallback"] > 1.0
