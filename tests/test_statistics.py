"""cfdr_amd.statistics on synthetic evaluations (no GPU).

What pins each function (rather than restating its formula):
* closed forms and hand-computed answers: weights (3:1 likelihoods give
  0.75 / 0.25), order statistics, CRPS of tiny forecasts, exact mixture
  variances, rank-bin cardinalities;
* brute force on small inputs: the O(M^2) pairwise CRPS definition, the
  conformal quantile as a sorted-index lookup, coverage over every
  permutation of three exchangeable labels;
* invariances: results independent of row order, output-ID chunking,
  additive likelihood offsets and translation of a forecast;
* the admission contract of the reference (rans-uq/python/cfdr/
  statistics.py:134-158, 208-232): which statuses keep predictions and
  which rows a posterior refuses.
The reference's random streams cannot be drawn here (its Python package is
not importable), so equality of streams with the reference is unpinned:
the streams are tested for determinism and for being keyed by every field.
"""
from __future__ import annotations

import itertools
import math
from types import SimpleNamespace

import numpy as np
import pytest

from cfdr_amd import engine
from cfdr_amd import statistics as st
from cfdr_amd.statistics import (Evaluation, EvaluationSet, Sample, conformal_quantile,
                                 empirical_crps, independent_uniform_design, posterior_weights,
                                 predictive_summary, randomized_ranks, rank_histograms,
                                 resample_posterior, sample_rng, split_conformal_intervals)

S = engine.EvaluationStatus


def ok(sample_id, loglik=0.0, predictions=(0.0,), theta=(0.31, 0.09)):
    return Evaluation(sample_id, theta, "Converged", predictions, loglik, 1, "synthetic", "")


def engine_row(status, predictions, loglik, iters=1):
    return SimpleNamespace(status=status, predictions=predictions, log_lik=loglik,
                           simple_iters=iters)


# ------------------------------------------------------------------ streams
def test_streams_are_deterministic_and_keyed_by_every_field():
    first = sample_rng(7, 3, "prior-design").random(3)
    assert np.array_equal(first, sample_rng(7, 3, "prior-design").random(3))
    others = [sample_rng(8, 3, "prior-design"), sample_rng(7, 4, "prior-design"),
              sample_rng(7, 3, "posterior-resample")]
    assert all(not np.array_equal(first, g.random(3)) for g in others)
    # the decimal fields are separated, so (1, 23) and (12, 3) are different keys
    assert sample_rng(1, 23, "x").random() != sample_rng(12, 3, "x").random()


def test_design_is_per_id_inside_the_box_and_leaves_the_global_rng_alone():
    np.random.seed(11)
    before = np.random.random()
    np.random.seed(11)
    a = independent_uniform_design([4, 9, 1], [0.28, 0.084], [0.34, 0.096], root_seed=20261010)
    assert np.random.random() == before
    b = independent_uniform_design([1, 4, 9], [0.28, 0.084], [0.34, 0.096], root_seed=20261010)
    assert [s.sample_id for s in a] == [4, 9, 1]
    assert {s.sample_id: s.theta for s in a} == {s.sample_id: s.theta for s in b}
    for s in a:
        assert 0.28 <= s.theta[0] < 0.34 and 0.084 <= s.theta[1] < 0.096
    draws = np.array([s.theta for s in independent_uniform_design(range(4000), [0.0], [1.0],
                                                                  root_seed=5)])
    assert abs(draws.mean() - 0.5) < 0.02 and abs(draws.var() - 1 / 12) < 0.005


# ---------------------------------------------------------------- admission
@pytest.mark.parametrize("status,preds,loglik,want_status,want_preds,want_ll,want_err", [
    (S.Converged, [2.0], -3.0, "Converged", (2.0,), -3.0, ""),
    (S.Converged, [], 0.0, "Unknown", (), None, "nonfinite or missing successful output"),
    (S.Converged, [np.nan], 0.0, "Unknown", None, None, "nonfinite or missing successful output"),
    (S.Converged, [np.inf], 0.0, "Unknown", None, None, "nonfinite or missing successful output"),
    (S.Converged, [2.0], np.nan, "Unknown", (2.0,), None, "nonfinite or missing successful output"),
    (S.Converged, [2.0], -np.inf, "Unknown", (2.0,), None,
     "nonfinite or missing successful output"),
    (S.Unconverged, [1.5], -2.0, "Unconverged", (1.5,), None, ""),
    (S.InvalidParameters, [np.nan], -1e30, "InvalidParameters", (), None, ""),
    (S.DivergenceDetected, [np.nan], -1e6, "DivergenceDetected", (), None, ""),
    (S.Unknown, [np.nan], -1e6, "Unknown", (), None, ""),
])
def test_engine_rows_become_evaluations_by_status(status, preds, loglik, want_status, want_preds,
                                                  want_ll, want_err):
    r = st._from_engine_row(Sample(9, (0.3, 0.09)), engine_row(status, preds, loglik, 7), "x")
    assert (r.sample_id, r.iterations, r.status, r.log_likelihood, r.error) == \
        (9, 7, want_status, want_ll, want_err)
    if want_preds is not None:
        assert r.predictions == want_preds
    assert r.admitted == (want_status == "Converged")


@pytest.mark.parametrize("status,error", [("Converged", "solver exception"),
                                          ("Unconverged", ""), ("DivergenceDetected", ""),
                                          ("InvalidParameters", ""), ("Unknown", "")])
def test_any_unadmitted_row_makes_the_posterior_undefined(status, error):
    bad = Evaluation(3, (0.3, 0.09), status, (0.0,), -1.0, 1, "synthetic", error)
    evs = EvaluationSet((ok(1), ok(2), bad))
    assert evs.failure_count == 1
    for action in (lambda: posterior_weights(evs), lambda: predictive_summary(evs, root_seed=1),
                   lambda: resample_posterior(evs, [0], root_seed=1)):
        with pytest.raises(ValueError, match="1 unresolved evaluations; posterior is undefined"):
            action()


def test_likelihood_sentinels_are_not_an_admission_rule():
    # a finite failure sentinel above every valid likelihood admits nothing,
    # and a valid likelihood far below it is still a weight of one
    np.testing.assert_array_equal(posterior_weights(EvaluationSet((ok(1, -2e30),))), [1.0])
    failed = Evaluation(2, (0.3, 0.09), "Unconverged", (0.0,), -1e6, 1, "synthetic", "")
    with pytest.raises(ValueError, match="unresolved"):
        posterior_weights(EvaluationSet((ok(1, -2e30), failed)))


# ------------------------------------------------------ importance posterior
def test_weights_closed_form_and_offset_invariance():
    evs = EvaluationSet((ok(5, math.log(3.0)), ok(2, 0.0, theta=(0.32, 0.095))))
    np.testing.assert_allclose(posterior_weights(evs), [0.75, 0.25], rtol=1e-15)
    shifted = EvaluationSet(tuple(ok(r.sample_id, r.log_likelihood - 1e4, theta=r.theta)
                                  for r in evs.records))
    np.testing.assert_allclose(posterior_weights(shifted), [0.75, 0.25], rtol=1e-12)
    rng = np.random.default_rng(3)
    ll = rng.normal(scale=40.0, size=50)
    w = posterior_weights(EvaluationSet(tuple(ok(i, float(v)) for i, v in enumerate(ll))))
    direct = np.exp(ll - ll.max())
    np.testing.assert_allclose(w, direct / direct.sum(), rtol=1e-12)
    assert math.isclose(float(w.sum()), 1.0, rel_tol=1e-14)


def test_resampling_ignores_input_order_and_output_chunking():
    evs = tuple(ok(i, float(-0.5 * (i - 3) ** 2), theta=(0.3 + 0.01 * i, 0.09)) for i in range(7))
    whole = resample_posterior(EvaluationSet(evs), range(100, 140), root_seed=17)
    parts = (resample_posterior(EvaluationSet(evs[::-1]), range(120, 140), root_seed=17) +
             resample_posterior(EvaluationSet(evs[3:] + evs[:3]), range(100, 120), root_seed=17))
    assert {s.sample_id: s.theta for s in whole} == {s.sample_id: s.theta for s in parts}


def test_resampling_frequencies_follow_the_weights():
    evs = EvaluationSet((ok(1, math.log(3.0)), ok(2, 0.0, theta=(0.32, 0.095))))
    draws = resample_posterior(evs, range(6000), root_seed=3)
    share = sum(s.theta == (0.31, 0.09) for s in draws) / 6000
    assert abs(share - 0.75) < 4 * math.sqrt(0.75 * 0.25 / 6000)


# --------------------------------------------------- predictive distribution
def test_mixture_variances_are_exact_and_labelled():
    same = EvaluationSet((ok(1, predictions=(4.0,)), ok(2, predictions=(4.0,))))
    bare = predictive_summary(same, root_seed=4)
    assert bare.latent_variance == bare.total_variance == (0.0,)
    assert bare.samples == ((4.0,), (4.0,))
    assert not (bare.includes_surrogate_noise or bare.includes_discrepancy or
                bare.includes_observation_noise)
    noisy = predictive_summary(same, root_seed=4, observation_sigma=(100.0,))
    assert noisy.observation_variance == noisy.total_variance == (10000.0,)
    assert noisy.includes_observation_noise and noisy.samples != bare.samples
    # predictions 0 and 2: population variance 1; 1 + 3 + 5 + 2^2 = 13
    two = EvaluationSet((ok(1, predictions=(0.0,)), ok(2, predictions=(2.0,))))
    full = predictive_summary(two, root_seed=4, observation_sigma=(2.0,),
                              surrogate_variance=(3.0,), discrepancy_variance=(5.0,))
    assert full.latent_mean == (1.0,) and full.latent_variance == (1.0,)
    assert full.total_variance == (13.0,)
    assert full.includes_surrogate_noise and full.includes_discrepancy


def test_each_records_noise_is_its_own_regardless_of_the_set():
    recs = tuple(ok(i, predictions=(float(i), -float(i))) for i in range(6))
    a = predictive_summary(EvaluationSet(recs), root_seed=9, observation_sigma=(0.5, 2.0))
    b = predictive_summary(EvaluationSet(recs[::-1]), root_seed=9, observation_sigma=(0.5, 2.0))
    assert dict(zip(a.sample_ids, a.samples)) == dict(zip(b.sample_ids, b.samples))
    assert a.latent_mean == b.latent_mean and a.total_variance == b.total_variance
    big = predictive_summary(EvaluationSet(tuple(ok(i, predictions=(1.0,)) for i in range(3000))),
                             root_seed=2, surrogate_variance=(4.0,))
    draws = np.asarray(big.samples)[:, 0]
    assert abs(draws.mean() - 1.0) < 0.15 and abs(draws.var() - 4.0) < 0.4


# -------------------------------------------------------------- conformal
@pytest.mark.parametrize("scores,alpha,expected",
                         [([1, 2, 3, 4], 0.5, 3), ([7], 0.1, math.inf), ([7], 0.5, 7),
                          ([1, 1, 2, 2], 0.5, 2)])
def test_conformal_order_statistic(scores, alpha, expected):
    assert conformal_quantile(scores, alpha) == expected


def test_conformal_quantile_is_the_kth_sorted_score():
    rng = np.random.default_rng(8)
    for n in (1, 2, 5, 19, 64):
        s = rng.exponential(size=n)
        for alpha in (0.05, 0.1, 0.3, 0.5, 0.9):
            k = math.ceil((n + 1) * (1 - alpha))
            want = math.inf if k > n else sorted(s)[k - 1]
            assert conformal_quantile(s, alpha) == want


def test_split_conformal_coverage_over_exchangeable_permutations():
    # constant predictor 0, one calibration label and one test label drawn
    # from {0, 1, 2} without replacement: at alpha 1/2 the interval is
    # 0 +/- |calibration label|, which covers the test label in 3 of 6 orders
    covered = 0
    for labels in itertools.permutations([0.0, 1.0, 2.0]):
        lo, hi = split_conformal_intervals([labels[0]], [0.0], [0.0], alpha=0.5)
        covered += int(lo[0] <= labels[1] <= hi[0])
    assert covered == 3


# ------------------------------------------------------------------- CRPS
def test_crps_closed_forms():
    # X uniform on {0, 2}, y = 1: E|X - y| = 1, E|X - X'| / 2 = 1/2
    assert empirical_crps([0.0, 2.0], 1.0) == 0.5
    assert empirical_crps([0.0], 1.0) == 1.0          # a point forecast: |x - y|
    assert empirical_crps([1e16, 1e16 + 2], 1e16) == 0.5
    assert empirical_crps([3.0, 3.0, 3.0], 3.0) == 0.0


def test_crps_equals_the_pairwise_definition_and_is_translation_invariant():
    rng = np.random.default_rng(5)
    for m in (2, 3, 9, 40):
        x, y = rng.normal(size=m), float(rng.normal())
        pairwise = np.mean(np.abs(x - y)) - 0.5 * np.mean(np.abs(x[:, None] - x[None, :]))
        assert empirical_crps(x, y) == pytest.approx(pairwise, rel=1e-12, abs=1e-15)
        assert empirical_crps(x + 1e3, y + 1e3) == pytest.approx(pairwise, rel=1e-9, abs=1e-12)


# ------------------------------------------------------------------- ranks
def test_tied_ranks_cover_every_position_and_follow_the_trial_id():
    ids = tuple(range(60))
    truth, draws = np.zeros((60, 2)), np.zeros((60, 5, 2))
    ranks = randomized_ranks(truth, draws, ids, root_seed=914)
    assert set(ranks[:, 0].tolist()) == set(range(6))
    assert not np.array_equal(ranks[:, 0], ranks[:, 1])
    flipped = randomized_ranks(truth[::-1], draws[::-1], ids[::-1], root_seed=914)
    np.testing.assert_array_equal(ranks, flipped[::-1])


def test_untied_ranks_count_the_draws_below():
    ranks = randomized_ranks([[2.0, 0.0]], [[[1.0, 1.0], [3.0, 2.0], [4.0, 3.0]]], [3],
                             root_seed=914)
    np.testing.assert_array_equal(ranks, [[1, 0]])


def test_rank_histograms_use_the_discrete_uniform_bin_sizes():
    # ranks 0..20 into 20 bins: one bin holds two ranks, the rest one each
    ranks = np.tile(np.arange(21), 100)[:, None]
    h = rank_histograms(ranks, n_draws=20, bins=20)
    assert sorted(h.expected_counts[0].tolist()) == [100.0] * 19 + [200.0]
    np.testing.assert_array_equal(h.counts, h.expected_counts)
    assert h.counts.sum() == ranks.size
    wide = rank_histograms(np.repeat(ranks, 3, axis=1), n_draws=20, bins=20)
    np.testing.assert_array_equal(wide.counts, np.repeat(h.counts, 3, axis=0))
    fewer = rank_histograms(ranks, n_draws=20, bins=64)  # at most n_draws + 1 bins
    assert fewer.counts.shape == (1, 21) and not hasattr(h, "pvalue")


# ------------------------------------------------------------ bad inputs
BAD_INPUTS = [
    (lambda: sample_rng(-1, 0, "x"), "root seed must be a nonnegative integer"),
    (lambda: sample_rng(1, True, "x"), "sample ID must be a nonnegative integer"),
    (lambda: sample_rng(1, 0, ""), "RNG purpose must be nonempty"),
    (lambda: Sample(1.5, (0.3,)), "sample ID must be a nonnegative integer"),
    (lambda: EvaluationSet((ok(2), ok(2))), "sample IDs must be unique"),
    (lambda: independent_uniform_design([2, 2], [0], [1], root_seed=1), "unique"),
    (lambda: independent_uniform_design([1], [1], [1], root_seed=1), "lower < upper"),
    (lambda: independent_uniform_design([1], [], [], root_seed=1), "lower < upper"),
    (lambda: independent_uniform_design([1], [0, 0], [1], root_seed=1), "lower < upper"),
    (lambda: independent_uniform_design([1], [-1e308], [1e308], root_seed=1), "lower < upper"),
    (lambda: Evaluation(1, (), "Done", (), None, 1, "x", ""), "unrecognized evaluation status"),
    (lambda: Evaluation(1, (), "Converged", (), None, 1, "", ""),
     "evaluation backend must be explicit"),
    (lambda: Evaluation(1, (), "Converged", (), None, -1, "x", ""),
     "iteration count must be a nonnegative integer"),
    (lambda: posterior_weights(EvaluationSet(())), "at least one evaluation is required"),
    (lambda: posterior_weights(EvaluationSet((ok(1, float("nan")),))),
     "successful log likelihoods must be finite"),
    (lambda: predictive_summary(EvaluationSet((ok(1, predictions=()),)), root_seed=1),
     "successful predictions must be finite equal-length nonempty vectors"),
    (lambda: predictive_summary(EvaluationSet((ok(1),)), root_seed=1, observation_sigma=(-1.0,)),
     "noise components must be finite nonnegative output-length vectors"),
    (lambda: predictive_summary(EvaluationSet((ok(1),)), root_seed=1,
                                surrogate_variance=(1.0, 1.0)),
     "noise components must be finite nonnegative output-length vectors"),
    (lambda: predictive_summary(EvaluationSet((ok(1),)), root_seed=1, observation_sigma=(1e200,)),
     "observation variance must be finite"),
    (lambda: predictive_summary(EvaluationSet((ok(1, predictions=(1e308,)),
                                               ok(2, predictions=(-1e308,)))), root_seed=1),
     "predictive samples and total variance must be finite"),
    (lambda: conformal_quantile([], 0.1), "conformal scores must be a finite nonempty vector"),
    (lambda: conformal_quantile([1, np.nan], 0.5), "finite nonempty vector"),
    (lambda: conformal_quantile([1], 0.0), "alpha must lie in"),
    (lambda: conformal_quantile([1], 1.0), "alpha must lie in"),
    (lambda: conformal_quantile([-1], 0.5), "scores must be nonnegative"),
    (lambda: split_conformal_intervals([1, 2], [1], [1], alpha=0.5),
     "calibration target and prediction shapes must match"),
    (lambda: split_conformal_intervals([[1]], [1], [1], alpha=0.5),
     "calibration targets must be a finite nonempty vector"),
    (lambda: empirical_crps([0, np.nan], 1), "forecast samples must be a finite nonempty vector"),
    (lambda: empirical_crps([0], np.inf), "truth must be finite"),
    (lambda: empirical_crps([-1e308, 1e308], 0.0), "CRPS arithmetic must remain finite"),
    (lambda: randomized_ranks([[0.0]], [[0.0]], [1], root_seed=1), "must align"),
    (lambda: randomized_ranks([[0.0]], [[[0.0]]], [1, 2], root_seed=1), "must align"),
    (lambda: randomized_ranks([[np.nan]], [[[0.0]]], [1], root_seed=1), "must align"),
    (lambda: rank_histograms([[0.5]], n_draws=2), "ranks must be a nonempty integer"),
    (lambda: rank_histograms([[3]], n_draws=2), "ranks must be a nonempty integer"),
    (lambda: rank_histograms([[0]], n_draws=0), "ranks must be a nonempty integer"),
    (lambda: rank_histograms([[0]], n_draws=2, bins=-1), "bin count must be a nonnegative integer"),
]


@pytest.mark.parametrize("call,message", BAD_INPUTS)
def test_invalid_inputs_raise_the_reference_messages(call, message):
    with pytest.raises(ValueError, match=message.replace("(", r"\(").replace("[", r"\[")):
        call()


def test_evaluators_check_their_backends_and_rows():
    with pytest.raises(TypeError, match="batch must be cfdr_amd.engine.ForwardBatch"):
        st.evaluate_batch([Sample(1, (0.3, 0.09))], object())
    with pytest.raises(ValueError, match="equal nonzero lengths"):
        st.evaluate_batch([Sample(1, (0.3, 0.09)), Sample(2, (0.3,))], object())
    with pytest.raises(ValueError, match="equal nonzero lengths"):
        st.evaluate_scalar([Sample(1, ())], object)
    with pytest.raises(TypeError, match="model_factory must construct"):
        st.evaluate_scalar([Sample(1, (0.3, 0.09))], object)
