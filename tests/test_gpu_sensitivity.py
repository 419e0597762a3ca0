"""Warm-FD total sensitivities on the GPU (rq_sensitivity_*): the 22
perturbed solves of one Jacobian in one launch from the device-resident base
state, against the oracle's restatement of ParameterSensitivity::
etaJacobianWarmFD (rans-uq/src/ParameterSensitivity.cpp:729-840) and the
contracts of the reference's own tests (rans-uq/tests/test_warm_fd_refinement.cpp,
rans-uq/tests/test_evaluation_safety.cpp:40-90)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py
from cfdr_amd import engine
from cfdr_amd.config import COEFF_DEFAULTS
from cfdr_amd.sensitivity import ParameterSensitivity
from fixtures import sens_channel, small_channel

pytestmark = pytest.mark.gpu

THETA = np.asarray(COEFF_DEFAULTS)


def _sens(prob):
    return ParameterSensitivity(prob.mesh, prob.obs, prob.bcs, prob.nu, prob.settings,
                                prob.u_init, prob.p_init, prob.k_init, prob.omega_init)


def _oracle(prob):
    return oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs, prob.nu,
                                   prob.init_vector())


def _iters_close(got, want):
    return np.all(np.abs(np.asarray(got) - np.asarray(want)) <= np.maximum(20, 0.005 * np.asarray(want)))


def test_warm_fd_matches_oracle_on_a_turbulent_channel():
    prob = small_channel()
    s = _sens(prob)
    assert s.solve_state(THETA) == engine.EvaluationStatus.Converged and s.has_state()
    g = s.eta_jacobian_warm_fd(THETA)
    o = _oracle(prob)
    st, state = oracle_py.solve_state(o, THETA)
    assert st == 0
    r = oracle_py.eta_jacobian_warm_fd(o, state, THETA)
    assert [int(v) for v in g.positive_status] == r["positive_status"].tolist()
    assert [int(v) for v in g.negative_status] == r["negative_status"].tolist()
    assert g.krylov_converged.tolist() == r["converged"].tolist()
    assert _iters_close(g.krylov_iters, r["iterations"])
    assert np.array_equal(g.finite_difference_steps, r["steps"])
    assert np.array_equal(g.perturbed_solve_tolerances, r["tolerances"])
    scale = np.maximum(np.max(np.abs(r["d_obs_d_theta"]), axis=0), 1e-3)
    assert np.all(np.abs(g.d_obs_d_theta - r["d_obs_d_theta"]) <= 1e-4 * scale)
    assert np.all(g.d_obs_d_theta[:, 10] == 0.0) and g.log_lik_gradient[10] == 0.0
    assert g.n_residual_evals == int(np.sum(g.krylov_iters))


def test_warm_fd_refinement_on_the_reference_problem():
    """test_warm_fd_refinement.cpp:48-66 on the GPU."""
    s = _sens(sens_channel())
    assert s.solve_state(THETA) == engine.EvaluationStatus.Converged
    steps = THETA * 2e-4
    coarse = s.eta_jacobian_warm_fd(THETA, 2e-4, 1e-9, 15000, 1e-10, steps)
    fine = s.eta_jacobian_warm_fd(THETA, 1e-4, 1e-9, 15000, 1e-11, steps / 2)
    assert coarse.krylov_converged.all() and fine.krylov_converged.all()
    a, b = coarse.d_obs_d_theta, fine.d_obs_d_theta
    assert np.all(np.abs(a - b) <= 1e-7 + 1e-3 * np.maximum(np.abs(a), np.abs(b)))


def test_warm_fd_contracts():
    """test_evaluation_safety.cpp:40-90: the base must be the converged
    coefficients; strict one-iteration perturbations fail as NaN columns with
    failing statuses; the default tolerance is min(tol, 0.01 h / |theta|);
    explicit steps and tolerances are taken as given; bad arguments raise."""
    s = _sens(sens_channel())
    with pytest.raises(ValueError, match="converged base solveState"):
        s.eta_jacobian_warm_fd(THETA)
    assert s.solve_state(np.full(11, 5.0)) == engine.EvaluationStatus.InvalidParameters
    assert not s.has_state()
    assert s.solve_state(THETA) == engine.EvaluationStatus.Converged
    other = THETA.copy()
    other[0] *= 1.01
    with pytest.raises(ValueError, match="coefficients of the converged base state"):
        s.eta_jacobian_warm_fd(other)
    failed = s.eta_jacobian_warm_fd(THETA, 5e-4, 1e-7, 1, 1e-14)
    assert not failed.krylov_converged.any()
    assert np.isnan(failed.d_obs_d_theta).all() and np.isnan(failed.log_lik_gradient).all()
    assert all(v != engine.EvaluationStatus.Converged
               for v in failed.positive_status + failed.negative_status)
    assert (failed.finite_difference_steps > 0).all()
    actual = s.eta_jacobian_warm_fd(THETA, 5e-4, 1e-7, 1)
    expected = np.minimum(1e-9, 0.01 * actual.finite_difference_steps / np.abs(THETA))
    assert np.all(np.abs(actual.perturbed_solve_tolerances - expected) <= 1e-20)
    explicit = s.eta_jacobian_warm_fd(THETA, 5e-4, 1e-7, 1, 1e-12, np.full(11, 2e-6))
    assert np.all(np.abs(explicit.finite_difference_steps - 2e-6) <= 1e-16)
    assert np.all(explicit.perturbed_solve_tolerances == 1e-12)
    with pytest.raises(ValueError, match="eleven positive finite"):
        s.eta_jacobian_warm_fd(THETA, 5e-4, 1e-7, 1, 0, [1e-5])
    with pytest.raises(ValueError, match="step/tolerance/iteration cap"):
        s.eta_jacobian_warm_fd(THETA, -1.0)
