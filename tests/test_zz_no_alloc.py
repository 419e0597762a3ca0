"""The allocation-free evaluation (rq_engine_evaluate_noalloc, no reference
counterpart): the same rows through the fixed device scratch region give the
results of rq_engine_evaluate, on two calls in a row (the second smaller, so
it runs over what the first left in the region), for an LDS-arena channel,
an HBM-arena hill with AMG spread over four workgroups per draw, and rows
the host settles without a launch; a call that would need more than the
region is refused with its size."""
from __future__ import annotations

import numpy as np
import pytest

from cfdr_amd import cases, engine
from fixtures import small_channel

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _batch(prob, **kw):
    return engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                               prob.u_init, prob.p_init, prob.k_init, prob.omega_init, **kw)


def _same(got, want):
    """Statuses, iteration counts and failure text exactly; observations and
    log-likelihoods within 1e-12 relative (the kernels are the same, so the
    bits are expected to match; the parity bar is 1e-6)."""
    assert [int(s) for s in got.statuses] == [int(s) for s in want.statuses]
    assert np.array_equal(got.iteration_counts, want.iteration_counts)
    assert np.array_equal(got.krylov_iterations, want.krylov_iterations)
    assert got.errors == want.errors
    assert np.array_equal(np.isnan(got.predictions), np.isnan(want.predictions))
    ok = ~np.isnan(want.predictions)
    rel = np.abs(got.predictions[ok] - want.predictions[ok]) / np.maximum(
        np.abs(want.predictions[ok]), 1e-300)
    assert rel.size == 0 or float(rel.max()) <= 1e-12
    ll = np.isfinite(want.log_likelihoods)
    assert np.array_equal(got.log_likelihoods[~ll], want.log_likelihoods[~ll])
    assert np.allclose(got.log_likelihoods[ll], want.log_likelihoods[ll], rtol=1e-12, atol=0)
    assert np.array_equal(got.convergence_mask, want.convergence_mask)


def test_noalloc_channel_matches_evaluate_twice():
    prob = small_channel()
    b = _batch(prob)
    draws = cases.channel_draws(6)
    # an out-of-bounds row and a nonfinite one are settled on the host
    first = np.vstack([draws[:4], [[0.9, 0.09]], [[np.nan, 0.09]]])
    second = draws[4:6]
    for th in (first, second):
        got = b.evaluate_noalloc(th)
        assert b.last_timing()["arena_in_lds"]
        _same(got, b.evaluate(th))
    b.close()


def test_noalloc_hill_multi_workgroup_matches_evaluate_twice():
    prob = cases.hill_problem(nx=30, ny=16, max_iterations=60)
    b = _batch(prob, grid_blocks=4)
    th = cases.hill_draws(3)
    for rows in (th[:3], th[:1]):
        got = b.evaluate_noalloc(rows)
        t = b.last_timing()
        assert not t["arena_in_lds"] and t["grid_blocks"] == 4
        _same(got, b.evaluate(rows))
    b.close()


def test_noalloc_only_host_rows_and_empty_call():
    prob = small_channel()
    b = _batch(prob)
    th = np.asarray([[0.9, 0.09], [0.31, 5.0]])
    got = b.evaluate_noalloc(th)
    _same(got, b.evaluate(th))
    assert b.evaluate_noalloc(np.zeros((0, 2))).statuses == []
    b.close()


def test_noalloc_refuses_rows_beyond_the_region():
    prob = cases.hill_problem(nx=120, ny=64, max_iterations=5)
    b = _batch(prob)
    arena = b.last_timing()["arena_bytes"]
    rows = SCRATCH // max(arena, 1) + 2
    th = np.tile(cases.hill_draws(1), (rows, 1))
    with pytest.raises(ValueError, match="bytes of device scratch"):
        b.evaluate_noalloc(th)
    # the engine still evaluates through the region after the refusal
    got = b.evaluate_noalloc(th[:1])
    assert [s.name for s in got.statuses] == ["Unconverged"]
    b.close()
