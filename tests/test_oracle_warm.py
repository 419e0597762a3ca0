"""The oracle's warm-started solves and its restatement of the warm-FD
Jacobian (CPU): warm starts are pinned bitwise against the reference's own
ForwardModel runs with its cache (fixtures ref_channel_warm.json and
ref_channel_c1_warm.json, rans-uq/src/ForwardModel.cpp:79-117); the warm-FD
Jacobian (rans-uq/src/ParameterSensitivity.cpp:729-840) against the contracts
of the reference's own tests (rans-uq/tests/test_warm_fd_refinement.cpp,
rans-uq/tests/test_evaluation_safety.cpp:40-90) and against what the
mathematics fixes (step refinement, the inert kappa column)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle_py
from cfdr_amd import cases
from cfdr_amd.config import COEFF_DEFAULTS
from fixtures import load, sens_channel, small_channel


def _oracle(prob):
    return oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs, prob.nu,
                                   prob.init_vector())


def _check(r, row):
    assert r["status"] == row["status"]
    assert r["iters"] == row["iters"]
    if row["status"] in (0, 1):
        assert np.array_equal(r["predictions"], np.asarray(row["predictions"]))
        assert r["loglik"] == row["loglik"]


def test_oracle_warm_sequence_equals_reference_forward_model_bitwise():
    d = load("ref_channel_warm.json")
    prob = small_channel()
    o = _oracle(prob)
    th = d["thetas"]
    r0 = o.evaluate(prob.coefficients(th[0]), save=True)
    _check(r0, d["rows"][0])
    # row 1 and row 3 start from row 0's state; row 4 from row 3's (the
    # reference's cache replaced row 0's entry by row 3's)
    r1 = o.evaluate(prob.coefficients(th[1]), warm=r0["state"])
    _check(r1, d["rows"][1])
    assert d["rows"][2]["status"] == 3 and not d["rows"][2]["used_warm_start"]
    r3 = o.evaluate(prob.coefficients(th[3]), warm=r0["state"], save=True)
    _check(r3, d["rows"][3])
    r4 = o.evaluate(prob.coefficients(th[4]), warm=r3["state"])
    _check(r4, d["rows"][4])


def test_oracle_warm_rows_on_the_config1_channel_equal_reference():
    d = load("ref_channel_c1_warm.json")
    prob = cases.channel_problem()
    o = _oracle(prob)
    r0 = o.evaluate(prob.coefficients(d["thetas"][0]), save=True)
    _check(r0, d["rows"][0])
    for th, row in zip(d["thetas"][1:], d["rows"][1:]):
        _check(o.evaluate(prob.coefficients(th), warm=r0["state"]), row)


def test_saved_state_restarts_to_the_same_fixed_point():
    prob = small_channel()
    o = _oracle(prob)
    c = prob.coefficients([0.31, 0.09])
    r0 = o.evaluate(c, save=True)
    again = o.evaluate(c, warm=r0["state"], save=True)
    assert again["status"] == 0 and again["iters"] < r0["iters"]
    assert np.max(np.abs(again["predictions"] - r0["predictions"]) /
                  np.abs(r0["predictions"])) < 1e-6
    assert again["state"][-2] == r0["state"][-2] == 1.0  # turbulence established


@pytest.fixture(scope="module")
def sens():
    prob = sens_channel()
    o = _oracle(prob)
    theta = np.asarray(COEFF_DEFAULTS)
    status, state = oracle_py.solve_state(o, theta)
    assert status == 0
    return o, state, theta


def _agree(a, b):
    return np.abs(a - b) <= 1e-7 + 1e-3 * np.maximum(np.abs(a), np.abs(b))


def test_warm_fd_refinement_on_the_reference_problem(sens):
    """test_warm_fd_refinement.cpp:48-66: steps theta * 2e-4 (tol 1e-10)
    against half of them (tol 1e-11) agree to 1e-7 + 1e-3 relative in every
    column, all converged."""
    o, state, theta = sens
    steps = theta * 2e-4
    coarse = oracle_py.eta_jacobian_warm_fd(o, state, theta, 2e-4, 1e-9, 15000, 1e-10, steps)
    fine = oracle_py.eta_jacobian_warm_fd(o, state, theta, 1e-4, 1e-9, 15000, 1e-11, steps / 2)
    assert coarse["converged"].all() and fine["converged"].all()
    assert _agree(coarse["d_obs_d_theta"], fine["d_obs_d_theta"]).all()
    assert np.array_equal(coarse["steps"], 0.5 * ((theta + steps) - (theta - steps)))


def test_warm_fd_on_a_turbulent_channel_refines_and_keeps_kappa_inert():
    """On the turbulent 1x32 channel the beta1, alpha1, betaStar and a1
    columns are large and refine like the reference's gate; kappa (index 10)
    enters no equation, so its column is exactly zero."""
    prob = small_channel()
    o = _oracle(prob)
    theta = np.asarray(COEFF_DEFAULTS)
    status, state = oracle_py.solve_state(o, theta)
    assert status == 0
    steps = theta * 2e-4
    coarse = oracle_py.eta_jacobian_warm_fd(o, state, theta, 2e-4, 1e-9, 40000, 1e-10, steps)
    fine = oracle_py.eta_jacobian_warm_fd(o, state, theta, 1e-4, 1e-9, 40000, 1e-11, steps / 2)
    assert coarse["converged"].all() and fine["converged"].all()
    a, b = coarse["d_obs_d_theta"], fine["d_obs_d_theta"]
    for j in (2, 3, 8, 9):
        assert np.max(np.abs(a[:, j])) > 1e-2
        assert _agree(a[:, j], b[:, j]).all(), j
    assert np.all(a[:, 10] == 0.0) and coarse["log_lik_gradient"][10] == 0.0


def test_warm_fd_failures_and_tolerances(sens):
    """test_evaluation_safety.cpp:55-82: one-iteration perturbations fail and
    leave NaN columns; the default tolerance is min(tol, 0.01 h / |theta|);
    explicit steps and tolerances are taken as given."""
    o, state, theta = sens
    failed = oracle_py.eta_jacobian_warm_fd(o, state, theta, 5e-4, 1e-7, 1, 1e-14)
    assert not failed["converged"].any()
    assert np.isnan(failed["d_obs_d_theta"]).all() and np.isnan(failed["log_lik_gradient"]).all()
    assert (failed["positive_status"] != 0).all() and (failed["negative_status"] != 0).all()
    assert (failed["steps"] > 0).all()
    actual = oracle_py.eta_jacobian_warm_fd(o, state, theta, 5e-4, 1e-7, 1)
    expected = np.minimum(1e-9, 0.01 * actual["steps"] / np.abs(theta))
    assert np.all(np.abs(actual["tolerances"] - expected) <= 1e-20)
    # the betaStar and a1 steps are capped at 1e-5 relative
    assert actual["steps"][8] == pytest.approx(1e-5 * theta[8], rel=1e-12)
    assert actual["steps"][0] == pytest.approx(5e-4 * theta[0], rel=1e-12)
    explicit = oracle_py.eta_jacobian_warm_fd(o, state, theta, 5e-4, 1e-7, 1, 1e-12,
                                              np.full(11, 2e-6))
    assert np.all(np.abs(explicit["steps"] - 2e-6) <= 1e-16)
    assert np.all(explicit["tolerances"] == 1e-12)
