"""The staged calibration driver (cfdr_amd.calibration) on analytic
likelihoods, without a GPU: every stage is checked against what the
mathematics fixes.

* The Laplace fit of an exact quadratic log-likelihood recovers its peak
  H^-1 g and covariance H^-1 to rounding, and refuses a saddle.
* The Student-t proposal's log density equals scipy's multivariate_t (an
  independent implementation), and its draws have the scale matrix's
  covariance dof / (dof - 2) * S.
* On a linear forward model with Gaussian noise the log-likelihood is an
  exact quadratic: the pattern search ends within its final step of the
  maximum, the Laplace centre is the maximum, and the importance posterior
  recovers the exact Gaussian posterior's mean and standard deviations.
* Any failed row stops the calibration (the reference refuses a posterior
  with unresolved probability mass).
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import stats as sps

from cfdr_amd.calibration import (Calibration, CalibrationSettings, laplace_fit,
                                  student_t_proposal)
from cfdr_amd.statistics import Evaluation, EvaluationSet


def test_laplace_fit_recovers_an_exact_quadratic():
    hess = np.array([[3.0, 0.5], [0.5, 2.0]])
    g = np.array([0.7, -0.4])
    x = np.array([(i, j) for i in range(-2, 3) for j in range(-2, 3)], dtype=float)
    loglik = 5.0 + x @ g - 0.5 * np.einsum("ni,ij,nj->n", x, hess, x)
    fit = laplace_fit(loglik, x)
    np.testing.assert_allclose(fit.peak, np.linalg.solve(hess, g), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(fit.covariance, np.linalg.inv(hess), rtol=1e-12, atol=1e-14)
    assert fit.residual_rms < 1e-12 and fit.loglik_range == pytest.approx(np.ptp(loglik))
    saddle = x[:, 0] ** 2 - x[:, 1] ** 2
    with pytest.raises(RuntimeError, match="not concave"):
        laplace_fit(saddle, x)


def test_student_t_proposal_density_and_spread():
    centre = np.array([0.32, 0.095])
    scale = np.array([[4e-6, 1e-7], [1e-7, 9e-7]])
    draws, log_q = student_t_proposal(centre, scale, range(5000), root_seed=7, dof=4)
    theta = np.array([d.theta for d in draws])
    want = sps.multivariate_t(loc=centre, shape=scale, df=4).logpdf(theta)
    np.testing.assert_allclose([log_q[d.sample_id] for d in draws], want, rtol=1e-10)
    cov = np.cov(theta.T)
    np.testing.assert_allclose(cov, 2.0 * scale, rtol=0.25, atol=2e-8)
    again, _ = student_t_proposal(centre, scale, [17, 3], root_seed=7, dof=4)
    assert again[0].theta == draws[17].theta and again[1].theta == draws[3].theta


class LinearChannel:
    """A stand-in forward model: predictions base * (1 + J (theta - mu)) at 30
    points, Gaussian noise sigma on the likelihood group, so log L is an
    exact quadratic with maximum mu and Hessian J_l^T diag(base/sigma)^2 J_l."""

    def __init__(self, mu=(0.322, 0.0952), fail_rows=()):
        rng = np.random.default_rng(1)
        self.mu = np.asarray(mu)
        self.base = np.linspace(2.0, 18.0, 30)
        self.jac = rng.normal(size=(30, 2)) * np.array([5.0, 12.0])
        self.group = np.arange(30) % 3
        self.sigma = 0.004 * self.base
        self.like = self.group == 0
        self.fail_rows = set(fail_rows)
        self.calls = 0

    def predictions(self, theta):
        return self.base * (1.0 + self.jac @ (np.asarray(theta) - self.mu))

    def loglik(self, theta):
        r = (self.predictions(theta) - self.base)[self.like] / self.sigma[self.like]
        return float(-0.5 * r @ r)

    def posterior_cov(self):
        j = (self.base[:, None] * self.jac / self.sigma[:, None])[self.like]
        return np.linalg.inv(j.T @ j)

    def __call__(self, samples):
        self.calls += 1
        out = []
        for s in samples:
            status = "Unconverged" if (self.calls, s.sample_id) in self.fail_rows else "Converged"
            out.append(Evaluation(s.sample_id, s.theta, status,
                                  tuple(self.predictions(s.theta).tolist()),
                                  self.loglik(s.theta) if status == "Converged" else None, 100,
                                  "analytic", ""))
        return EvaluationSet(tuple(out))


def test_staged_calibration_finds_the_peak_and_its_laplace_fit():
    model = LinearChannel()
    res = Calibration(model).run(model.base, model.sigma, model.group)
    step = res.final_step
    assert np.all(np.abs(res.optimum - model.mu) <= step + 1e-12)
    # an exact quadratic: the Laplace centre and covariance are the posterior's
    np.testing.assert_allclose(res.center, model.mu, rtol=0, atol=1e-9)
    cov = model.posterior_cov()
    np.testing.assert_allclose(res.scale, cov, rtol=1e-6)
    sd = np.sqrt(np.diag(cov))
    for post in (res.main, res.replicate):
        # 64 Student-t(4) draws: loose bounds (the proposal has no fourth moment)
        assert np.all(np.abs(post.theta_mean - model.mu) < 0.5 * sd)
        assert 20 < post.effective_sample_size <= 64 and post.distinct_draws > 10
    polls = len(res.search_path) - 1 + CalibrationSettings().search_halvings
    assert res.solves == 24 + 8 * polls + 25 + 2 * 64
    assert res.launches == 1 + polls + 1 + 2
    assert res.coverage["held_out"] == 10
    assert 0 <= res.coverage["recalibrated_covered"] <= 10


def test_importance_posterior_converges_to_the_gaussian_posterior():
    model = LinearChannel()
    res = Calibration(model, CalibrationSettings(proposal_draws=3000)).run(
        model.base, model.sigma, model.group)
    sd = np.sqrt(np.diag(model.posterior_cov()))
    for post in (res.main, res.replicate):
        assert np.all(np.abs(post.theta_mean - model.mu) < 0.08 * sd)
        np.testing.assert_allclose(post.theta_sd, sd, rtol=0.05)


def test_a_failed_row_stops_the_calibration():
    model = LinearChannel(fail_rows={(2, 3)})  # row 3 of the first pattern-search poll
    with pytest.raises(RuntimeError, match="1 solves failed; the posterior is undefined"):
        Calibration(model).run(model.base, model.sigma, model.group)
