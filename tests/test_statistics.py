"""The statistics layer on synthetic evaluations (no GPU): the reference's
hand-calculated answers and contracts (rans-uq/python/tests/test_statistics.py)
restated against cfdr_amd.statistics."""
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


def record(sample_id, *, status="Converged", likelihood=0.0, prediction=(0.0,),
           theta=(0.31, 0.09)):
    return Evaluation(sample_id, theta, status, prediction, likelihood, 1, "analytic", "")


def test_importance_weights_and_resampling_use_ids():
    records = (record(5, likelihood=math.log(3)), record(2, theta=(0.32, 0.095)))
    np.testing.assert_allclose(posterior_weights(EvaluationSet(records)), [0.75, 0.25])
    a = resample_posterior(EvaluationSet(records), [10, 11, 12], root_seed=17)
    b = resample_posterior(EvaluationSet(records[::-1]), [12, 10, 11], root_seed=17)
    assert {s.sample_id: s.theta for s in a} == {s.sample_id: s.theta for s in b}


def test_resampling_frequencies_follow_the_weights():
    records = (record(1, likelihood=math.log(3)), record(2, theta=(0.32, 0.095)))
    draws = resample_posterior(EvaluationSet(records), range(4000), root_seed=3)
    share = sum(1 for s in draws if s.theta == (0.31, 0.09)) / 4000
    assert abs(share - 0.75) < 0.03


def test_likelihood_magnitude_never_admits_failure_or_rejects_valid_value():
    valid = EvaluationSet((record(1, likelihood=-2e30),))
    np.testing.assert_array_equal(posterior_weights(valid), [1.0])
    failed = record(2, status="Unconverged", likelihood=-1e30)
    with pytest.raises(ValueError, match="unresolved"):
        posterior_weights(EvaluationSet((valid.records[0], failed)))
    with pytest.raises(ValueError, match="finite"):
        posterior_weights(EvaluationSet((record(1, likelihood=float("nan")),)))


def test_duplicate_sample_ids_are_rejected():
    with pytest.raises(ValueError, match="unique"):
        EvaluationSet((record(2), record(2)))
    with pytest.raises(ValueError, match="unique"):
        independent_uniform_design([2, 2], [0], [1], root_seed=1)


def test_design_streams_do_not_use_global_rng_or_row_order():
    np.random.seed(55)
    expected = np.random.random()
    np.random.seed(55)
    first = independent_uniform_design([8, 2, 5], [0, 2], [1, 3], root_seed=812)
    assert tuple(s.sample_id for s in first) == (8, 2, 5)
    assert np.random.random() == expected
    np.random.random(17)
    second = independent_uniform_design([5, 8, 2], [0, 2], [1, 3], root_seed=812)
    assert {s.sample_id: s.theta for s in first} == {s.sample_id: s.theta for s in second}
    assert sample_rng(812, 8, "design").random() != sample_rng(812, 8, "prediction").random()
    assert all(0 <= s.theta[0] < 1 and 2 <= s.theta[1] < 3 for s in first)


def test_streams_are_keyed_by_every_field():
    base = sample_rng(1, 2, "x").random(4)
    assert np.array_equal(base, sample_rng(1, 2, "x").random(4))
    for other in (sample_rng(2, 2, "x"), sample_rng(1, 3, "x"), sample_rng(1, 2, "y")):
        assert not np.array_equal(base, other.random(4))
    # decimal fields are separated: (1, 23) and (12, 3) differ
    assert sample_rng(1, 23, "x").random() != sample_rng(12, 3, "x").random()


@pytest.mark.parametrize("scores,alpha,expected",
                         [([1, 2, 3, 4], 0.5, 3), ([7], 0.1, math.inf), ([7], 0.5, 7),
                          ([1, 1, 2, 2], 0.5, 2)])
def test_conformal_exact_order_statistic(scores, alpha, expected):
    assert conformal_quantile(scores, alpha) == expected


@pytest.mark.parametrize("scores,alpha", [([], 0.1), ([1, np.nan], 0.5), ([1, np.inf], 0.5),
                                          ([1], 0), ([1], 1), ([1], np.nan), ([-1], 0.5)])
def test_conformal_rejects_invalid_domains(scores, alpha):
    with pytest.raises(ValueError):
        conformal_quantile(scores, alpha)


def test_ordinary_split_coverage_on_exact_exchangeable_sweep_labels():
    # fixed predictor 0, one calibration point, one test point: over the six
    # permutations of {0, 1, 2}, 3 of 6 tests fall inside
    covered = 0
    for labels in itertools.permutations([0.0, 1.0, 2.0]):
        lo, hi = split_conformal_intervals([labels[0]], [0], [0], alpha=0.5)
        covered += lo[0] <= labels[1] <= hi[0]
    assert covered == 3


def test_rank_histogram_counts_discrete_cardinality_per_parameter():
    ranks = np.tile(np.arange(21), 100)[:, None]
    d = rank_histograms(ranks, n_draws=20, bins=20)
    np.testing.assert_array_equal(d.counts, d.expected_counts)
    assert sorted(d.expected_counts[0]) == [100] * 19 + [200]
    dup = rank_histograms(np.repeat(ranks, 2, axis=1), n_draws=20, bins=20)
    np.testing.assert_array_equal(dup.counts, np.repeat(d.counts, 2, axis=0))
    assert not hasattr(d, "pvalue")


def test_rank_ties_are_randomized_by_trial_id_and_parameter():
    ids = tuple(range(60))
    truth, draws = np.zeros((60, 2)), np.zeros((60, 5, 2))
    ranks = randomized_ranks(truth, draws, ids, root_seed=914)
    assert set(ranks[:, 0]) == set(range(6))
    assert not np.array_equal(ranks[:, 0], ranks[:, 1])
    reverse = randomized_ranks(truth[::-1], draws[::-1], ids[::-1], root_seed=914)
    np.testing.assert_array_equal(ranks, reverse[::-1])
    untied = randomized_ranks([[2, 0]], [[[1, 1], [3, 2], [4, 3]]], [3], root_seed=914)
    np.testing.assert_array_equal(untied, [[1, 0]])


def test_prediction_labels_and_variance_components():
    det = EvaluationSet((record(1, prediction=(4.0,)), record(2, prediction=(4.0,))))
    latent = predictive_summary(det, root_seed=4)
    assert latent.latent_variance == latent.total_variance == (0.0,)
    assert not (latent.includes_surrogate_noise or latent.includes_discrepancy or
                latent.includes_observation_noise)
    observed = predictive_summary(det, root_seed=4, observation_sigma=(100.0,))
    assert observed.observation_variance == observed.total_variance == (10000.0,)
    assert observed.includes_observation_noise and observed.samples != latent.samples
    mixed = EvaluationSet((record(1, prediction=(0.0,)), record(2, prediction=(2.0,))))
    c = predictive_summary(mixed, root_seed=4, observation_sigma=(2.0,),
                           surrogate_variance=(3.0,), discrepancy_variance=(5.0,))
    assert c.latent_mean == (1.0,) and c.latent_variance == (1.0,)
    assert c.total_variance == (13.0,)
    assert c.includes_surrogate_noise and c.includes_discrepancy


def test_empirical_crps_has_finite_forecast_denominator():
    # E|X-1| = 1 and 0.5 E|X-X'| = 0.5 for X uniform on {0, 2}
    assert empirical_crps([0, 2], 1) == 0.5
    assert empirical_crps([0], 1) == 1
    assert empirical_crps([1e16, 1e16 + 2], 1e16) == 0.5
    with pytest.raises(ValueError):
        empirical_crps([0, np.nan], 1)


def test_empirical_crps_equals_pairwise_definition():
    rng = np.random.default_rng(5)
    x, y = rng.normal(size=9), 0.3
    direct = np.mean(np.abs(x - y)) - 0.5 * np.mean(np.abs(x[:, None] - x[None, :]))
    assert empirical_crps(x, y) == pytest.approx(direct, rel=1e-12)


@pytest.mark.parametrize("error", ["solver exception", "invalid native diagnostic"])
def test_success_status_with_error_is_unresolved(error):
    ev = Evaluation(3, (0.31, 0.09), "Converged", (2.0,), 0.0, 1, "analytic", error)
    evs = EvaluationSet((ev,))
    assert not ev.admitted and evs.failure_count == 1
    for action in (lambda: posterior_weights(evs), lambda: predictive_summary(evs, root_seed=17)):
        with pytest.raises(ValueError, match="unresolved"):
            action()


def _native(status, prediction, likelihood, iters=1):
    return SimpleNamespace(status=status, predictions=prediction, log_lik=likelihood,
                           simple_iters=iters)


@pytest.mark.parametrize("prediction,likelihood",
                         [([], 0.0), ([np.nan], 0.0), ([np.inf], 0.0), ([2.0], np.nan),
                          ([2.0], np.inf)])
def test_admission_rejects_invalid_native_success(prediction, likelihood):
    r = st._record(Sample(9, (0.31, 0.09)),
                   _native(engine.EvaluationStatus.Converged, prediction, likelihood), "analytic")
    assert r.sample_id == 9 and r.status == "Unknown" and r.log_likelihood is None
    assert r.error == "nonfinite or missing successful output"
    with pytest.raises(ValueError, match="unresolved"):
        posterior_weights(EvaluationSet((r,)))


def test_admission_keeps_a_valid_native_success_and_drops_failed_predictions():
    ok = st._record(Sample(9, (0.31, 0.09)),
                    _native(engine.EvaluationStatus.Converged, [2.0], -3.0, 4), "analytic")
    assert ok.predictions == (2.0,) and ok.log_likelihood == -3.0 and ok.iterations == 4
    np.testing.assert_array_equal(posterior_weights(EvaluationSet((ok,))), [1.0])
    for status in (engine.EvaluationStatus.InvalidParameters, engine.EvaluationStatus.Unknown):
        bad = st._record(Sample(1, (0.31, 0.09)), _native(status, [np.nan], -1e30), "analytic")
        assert bad.predictions == () and bad.log_likelihood is None
    capped = st._record(Sample(2, (0.31, 0.09)),
                        _native(engine.EvaluationStatus.Unconverged, [1.5], -2.0), "analytic")
    assert capped.predictions == (1.5,) and capped.log_likelihood is None
    assert not capped.admitted


@pytest.mark.parametrize("bad", [-1, True, 1.5])
def test_statistics_rejects_invalid_identifiers(bad):
    with pytest.raises(ValueError, match="nonnegative integer"):
        Sample(bad, (0.31, 0.09))
    with pytest.raises(ValueError, match="nonnegative integer"):
        sample_rng(bad, 3, "design")


def test_rng_requires_an_explicit_purpose():
    with pytest.raises(ValueError, match="purpose must be nonempty"):
        sample_rng(1, 2, "")


@pytest.mark.parametrize("lower,upper", [([], []), ([1], [0]), ([0], [np.inf]), ([0], [1, 2])])
def test_design_rejects_invalid_bounds(lower, upper):
    with pytest.raises(ValueError, match="uniform bounds"):
        independent_uniform_design([3], lower, upper, root_seed=1)


def test_evaluation_requires_known_status_and_backend():
    with pytest.raises(ValueError, match="unrecognized evaluation status"):
        record(1, status="success")
    with pytest.raises(ValueError, match="backend must be explicit"):
        Evaluation(1, (0.31, 0.09), "Converged", (1.0,), 0.0, 1, "", "")


@pytest.mark.parametrize("samples", [(Sample(1, ()),),
                                     (Sample(1, (0.3,)), Sample(2, (0.3, 0.09)))])
def test_evaluation_rejects_incompatible_parameter_vectors(samples):
    with pytest.raises(ValueError, match="equal nonzero lengths"):
        st.evaluate_scalar(samples, object)


def test_evaluation_requires_engine_backends_and_positive_workers():
    with pytest.raises(TypeError, match="model_factory must construct"):
        st.evaluate_scalar((Sample(1, (0.31, 0.09)),), object)
    with pytest.raises(TypeError, match="batch must be"):
        st.evaluate_batch((), object())


def test_posterior_requires_nonempty_resolved_predictions():
    with pytest.raises(ValueError, match="at least one evaluation"):
        posterior_weights(EvaluationSet(()))
    for prediction in ((), (np.nan,)):
        with pytest.raises(ValueError, match="successful predictions must be finite"):
            predictive_summary(EvaluationSet((record(1, prediction=prediction),)), root_seed=1)


@pytest.mark.parametrize("component",
                         ["observation_sigma", "surrogate_variance", "discrepancy_variance"])
@pytest.mark.parametrize("value", [(-1.0,), (np.inf,), (0.0, 1.0)])
def test_predictive_noise_requires_finite_nonnegative_output_dimensions(component, value):
    with pytest.raises(ValueError, match="noise components must be finite nonnegative"):
        predictive_summary(EvaluationSet((record(1),)), root_seed=1, **{component: value})


def test_predictive_variance_overflow_is_an_explicit_failure():
    evs = EvaluationSet((record(1),))
    with np.errstate(over="ignore", invalid="ignore"):
        with pytest.raises(ValueError, match="observation variance must be finite"):
            predictive_summary(evs, root_seed=1, observation_sigma=(1e308,))
        with pytest.raises(ValueError, match="predictive samples and total variance must be finite"):
            predictive_summary(evs, root_seed=1, surrogate_variance=(1e308,),
                               discrepancy_variance=(1e308,))


def test_conformal_calibration_vectors_must_align():
    with pytest.raises(ValueError, match="target and prediction shapes must match"):
        split_conformal_intervals([1, 2], [0], [0], alpha=0.5)


def test_crps_rejects_nonfinite_truth_and_arithmetic_overflow():
    with pytest.raises(ValueError, match="truth must be finite"):
        empirical_crps([0], np.nan)
    with np.errstate(over="ignore", invalid="ignore"), \
            pytest.raises(ValueError, match="CRPS arithmetic must remain finite"):
        empirical_crps([-1e308, 1e308], -1e308)


@pytest.mark.parametrize("truth,draws,ids", [([], [], []), ([[0]], [[[0]]], []),
                                             ([[0]], [[[np.nan]]], [1]), ([[0]], [[]], [1])])
def test_rank_inputs_require_aligned_finite_trials(truth, draws, ids):
    with pytest.raises(ValueError, match="must align"):
        randomized_ranks(truth, draws, ids, root_seed=1)


@pytest.mark.parametrize("ranks", [[], [[-1]], [[3]], [[1.5]]])
def test_rank_histograms_reject_invalid_discrete_ranks(ranks):
    with pytest.raises(ValueError, match="nonempty integer"):
        rank_histograms(ranks, n_draws=2)
