"""The statistics layer over the GPU engine (reference contracts of
rans-uq/python/tests/test_statistics.py:43-100): failed rows keep their IDs
and statuses on both the scalar (ForwardModel) and batch (ForwardBatch)
paths, converged rows agree between the two, and the posterior is refused
while any row is unresolved."""
from __future__ import annotations

import numpy as np
import pytest

from cfdr_amd import engine
from cfdr_amd import statistics as st
from cfdr_amd.model import ForwardModel
from cfdr_amd.statistics import Sample
from fixtures import small_channel

pytestmark = pytest.mark.gpu


def _inputs(max_iterations):
    p = small_channel(max_iterations)
    return (p.mesh, p.params, p.obs, p.bcs, p.nu, p.settings, p.u_init, p.p_init, p.k_init,
            p.omega_init)


def test_native_admission_keeps_failed_ids_and_aligned_statuses():
    samples = (Sample(8, (0.31, 0.09)), Sample(3, (-1.0, 0.09)), Sample(9, (0.3, 0.08)))
    inputs = _inputs(0)
    scalar = st.evaluate_scalar(samples, lambda: ForwardModel(*inputs))
    one = st.evaluate_batch(samples, engine.ForwardBatch(*inputs), workers=1)
    two = st.evaluate_batch(samples[::-1], engine.ForwardBatch(*inputs), workers=2)
    assert scalar.failure_count == one.failure_count == two.failure_count == 3
    assert tuple(r.sample_id for r in scalar.records) == (8, 3, 9)
    assert tuple(r.status for r in scalar.records) == ("Unconverged", "InvalidParameters",
                                                       "Unconverged")
    for a, b, c in zip(scalar.records, one.records, two.records[::-1]):
        assert a.sample_id == b.sample_id == c.sample_id
        assert a.status == b.status == c.status
        assert a.log_likelihood is b.log_likelihood is c.log_likelihood is None
        np.testing.assert_array_equal(a.predictions, b.predictions)
        np.testing.assert_array_equal(b.predictions, c.predictions)
        assert a.backend == st.BACKEND_MODEL and b.backend == st.BACKEND_BATCH
    with pytest.raises(ValueError, match="unresolved"):
        st.posterior_weights(one)


def test_converged_rows_match_between_scalar_and_batch_and_give_a_posterior():
    inputs = _inputs(40000)
    samples = (Sample(8, (0.31, 0.09)), Sample(3, (0.32, 0.095)))
    scalar = st.evaluate_scalar(samples, lambda: ForwardModel(*inputs))
    batch = st.evaluate_batch(samples[::-1], engine.ForwardBatch(*inputs), workers=2)
    assert scalar.failure_count == batch.failure_count == 0
    for a, b in zip(scalar.records, batch.records[::-1]):
        assert a.status == b.status == "Converged"
        assert a.log_likelihood == b.log_likelihood
        assert a.iterations == b.iterations
        assert a.predictions == b.predictions
    w = st.posterior_weights(batch)
    assert np.all(np.isfinite(w)) and abs(w.sum() - 1.0) < 1e-12
    a = st.resample_posterior(scalar, [1, 2, 3], root_seed=5)
    b = st.resample_posterior(batch, [3, 1, 2], root_seed=5)
    assert {s.sample_id: s.theta for s in a} == {s.sample_id: s.theta for s in b}
    summary = st.predictive_summary(batch, root_seed=9, observation_sigma=[0.1] * 30)
    assert len(summary.total_variance) == 30
