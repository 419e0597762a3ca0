"""The single-row model with its device-resident warm-start cache (rq_model_*)
on the GPU: a warm-started sequence against the reference's own ForwardModel
run with its cache (rans-uq/src/ForwardModel.cpp:79-117, fixtures
ref_channel_warm.json and ref_channel_c1_warm.json), cold rows equal to the
batch path, the byte budget, and the one-launch batch mode against single
warm-started rows."""
from __future__ import annotations

import numpy as np
import pytest

from cfdr_amd import cases, engine
from cfdr_amd.model import ForwardModel
from fixtures import load, max_rel, small_channel

pytestmark = pytest.mark.gpu

REL_OBS = 1e-6


def _model(prob, **kw):
    return ForwardModel(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                        prob.u_init, prob.p_init, prob.k_init, prob.omega_init, **kw)


def test_warm_sequence_matches_reference_forward_model():
    d = load("ref_channel_warm.json")
    m = _model(small_channel())
    res, used = m.evaluate_sequence(np.asarray(d["thetas"]))
    for i, row in enumerate(d["rows"]):
        assert bool(used[i]) == row["used_warm_start"], i
        assert res.statuses[i] == row["status"], (i, res.errors[i])
        if row["status"] == 0:
            # converged solves: the contraction measure may cross tol a monitor
            # interval or two apart (parallel reductions)
            assert abs(int(res.iteration_counts[i]) - row["iters"]) <= max(20, int(0.005 * row["iters"]))
            assert max_rel(res.predictions[i], row["predictions"]) <= REL_OBS
            assert res.log_likelihoods[i] == pytest.approx(row["loglik"], rel=REL_OBS)
        else:
            assert res.iteration_counts[i] == row["iters"]
            assert np.all(np.isnan(res.predictions[i]))
            assert res.log_likelihoods[i] == row["loglik"]
    # converged rows are cached; the repeated row replaced its entry
    assert m.cache().size() == 2
    assert m.cache().retained_bytes() == 2 * (2 * 8 + (32 + 2) * 24 + 8 * (32 + 2) * 8)


def test_config1_warm_rows_are_classified_as_the_reference_does():
    """On the 1x128 channel (tol 1e-8) the reference classifies warm starts
    from the converged base as diverged (iterations 5, 332, 5): failed solves
    match exactly."""
    d = load("ref_channel_c1_warm.json")
    m = _model(cases.channel_problem())
    res, used = m.evaluate_sequence(np.asarray(d["thetas"]))
    for i, row in enumerate(d["rows"]):
        assert bool(used[i]) == row["used_warm_start"], i
        assert res.statuses[i] == row["status"], (i, res.errors[i])
        if row["status"] == 0:
            assert abs(int(res.iteration_counts[i]) - row["iters"]) <= max(20, int(0.005 * row["iters"]))
            assert max_rel(res.predictions[i], row["predictions"]) <= REL_OBS
        else:
            assert res.iteration_counts[i] == row["iters"], i
            assert res.log_likelihoods[i] == row["loglik"]
    assert m.cache().size() == 1


def test_cold_row_equals_batch_row_and_warm_restart_is_short():
    prob = small_channel()
    theta = np.asarray([0.31, 0.09])
    batch = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu,
                                prob.settings, prob.u_init, prob.p_init, prob.k_init,
                                prob.omega_init).evaluate(theta[None, :])
    m = _model(prob)
    first = m.evaluate_with_diagnostics(theta)
    assert not first.used_warm_start
    assert first.result.status == batch.statuses[0]
    assert first.result.simple_iters == batch.iteration_counts[0]
    assert np.array_equal(first.result.predictions, batch.predictions[0])
    again = m.evaluate_with_diagnostics(theta)
    assert again.used_warm_start and again.result.status == engine.EvaluationStatus.Converged
    assert again.result.simple_iters < first.result.simple_iters // 4
    assert max_rel(again.result.predictions, first.result.predictions) <= REL_OBS
    assert m.penalized_log_likelihood(theta) == pytest.approx(first.result.log_lik, rel=REL_OBS)


def test_zero_byte_budget_never_warm_starts():
    prob = small_channel()
    m = _model(prob, cache_byte_budget=0)
    theta = np.asarray([[0.31, 0.09], [0.31, 0.09]])
    res, used = m.evaluate_sequence(theta)
    assert not used.any() and m.cache().size() == 0
    assert np.array_equal(res.predictions[0], res.predictions[1])
    assert res.iteration_counts[0] == res.iteration_counts[1]
    # raising the budget afterwards enables the cache
    m.cache().set_byte_budget(1 << 20)
    m.evaluate(theta[0])
    assert m.cache().size() == 1 and m.evaluate_with_diagnostics(theta[0]).used_warm_start
    m.cache().clear()
    assert m.cache().size() == 0 and not m.evaluate_with_diagnostics(theta[0]).used_warm_start


def test_batch_mode_rows_equal_single_warm_rows():
    """evaluate_batch: every row warm-starts from the cache as it stood at
    the call, so each row equals the same row evaluated alone after the same
    cache history."""
    prob = small_channel()
    base = np.asarray([0.31, 0.09])
    stencil = np.asarray([[0.315, 0.09], [0.305, 0.09], [0.31, 0.0915], [0.31, 0.0885],
                          [0.9, 0.09]])
    mb = _model(prob)
    mb.evaluate(base)
    res, used = mb.evaluate_batch(stencil)
    assert used.tolist() == [True, True, True, True, False]
    assert res.statuses[4] == engine.EvaluationStatus.InvalidParameters
    assert mb.cache().size() == 1 + int(np.sum(res.convergence_mask))
    for i in range(4):
        ms = _model(prob)
        ms.evaluate(base)
        one = ms.evaluate_with_diagnostics(stencil[i])
        assert one.used_warm_start
        assert one.result.status == res.statuses[i]
        assert one.result.simple_iters == res.iteration_counts[i]
        assert np.array_equal(one.result.predictions, res.predictions[i])
