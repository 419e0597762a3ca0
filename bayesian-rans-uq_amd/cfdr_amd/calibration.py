"""The calibration sequence of the reference's coverage study, run on the GPU
engine (docs/figures/calibration_coverage.py:81-261 in the reference).

Because the likelihood of a channel calibration is sharply peaked, the
posterior is found in stages, each a batch of independent forward solves:

1. a seeded uniform design over a box of (a1, beta*) locates the peak
   region (24 solves; these solves never enter the posterior);
2. a pattern search moves a 3x3 stencil to its best point and halves the
   step whenever the centre wins, three times (8 solves per poll);
3. a least-squares quadratic of the log-likelihood on a 5x5 grid of the
   final step around the optimum gives a Laplace approximation: centre and
   covariance (25 solves);
4. importance sampling from a Student-t proposal with that centre and scale
   (64 solves): weights are likelihood over proposal density under a flat
   prior, 400 posterior draws are resampled, and their predictive
   distribution adds the observation noise; a second independent proposal
   sample checks the result (64 more solves);
5. split-conformal intervals on the log scale recalibrate the predictive
   ranges with a third of the points, and the held-out third measures
   coverage.

Every stage refuses failed solves exactly as the reference does: any row
that is not Converged stops the calibration (a posterior with unresolved
probability mass is undefined). Each batch is one GPU launch, so the
sequence costs about fifteen launches rather than 241 serial solves.
"""
from __future__ import annotations

import dataclasses
import math
from collections.abc import Callable, Sequence
from dataclasses import dataclass, field

import numpy as np

from . import statistics as stats
from .config import FlowBoundaryConditions, InferenceParameterSet, ObservationOperator, Vec3
from .engine import ForwardBatch
from .statistics import EvaluationSet, Sample

Evaluator = Callable[[tuple], EvaluationSet]


@dataclass(frozen=True)
class CalibrationSettings:
    """The study's constants (calibration_coverage.py:47-54)."""
    seed: int = 20261010
    alpha: float = 0.1
    stage1_box: tuple = ((0.28, 0.084), (0.34, 0.096))
    stage1_draws: int = 24
    search_step: tuple = (0.004, 0.001)
    search_halvings: int = 3
    grid_half_width: int = 2          # a (2k+1) x (2k+1) Laplace grid
    proposal_draws: int = 64
    proposal_dof: int = 4
    posterior_draws: int = 400


# --------------------------------------------------------------- stage 3
@dataclass(frozen=True)
class LaplaceFit:
    peak: np.ndarray        # in the grid's scaled coordinates
    covariance: np.ndarray  # in the grid's scaled coordinates
    residual_rms: float
    loglik_range: float


def laplace_fit(loglik: np.ndarray, x: np.ndarray) -> LaplaceFit:
    """Least-squares fit of log L = c + g.x - x.H.x / 2 over points x (n x 2):
    the Gaussian approximation has mean H^-1 g and covariance H^-1. H must
    be positive definite (the log-likelihood concave on the grid)."""
    x = np.asarray(x, dtype=float)
    loglik = np.asarray(loglik, dtype=float)
    u, v = x[:, 0], x[:, 1]
    basis = np.column_stack([np.ones_like(u), u, v, u * u, u * v, v * v])
    c, *_ = np.linalg.lstsq(basis, loglik, rcond=None)
    hessian = -np.array([[2.0 * c[3], c[4]], [c[4], 2.0 * c[5]]])
    if np.linalg.eigvalsh(hessian).min() <= 0.0:
        raise RuntimeError("the log-likelihood is not concave on the fitting grid")
    residual = loglik - basis @ c
    return LaplaceFit(np.linalg.solve(hessian, c[1:3]), np.linalg.inv(hessian),
                      float(np.sqrt(np.mean(residual * residual))), float(np.ptp(loglik)))


# --------------------------------------------------------------- stage 4
def student_t_proposal(center, scale, sample_ids: Sequence[int], *, root_seed: int,
                       dof: int) -> tuple:
    """Multivariate Student-t draws (location center, scale matrix scale,
    dof degrees of freedom), each from its own "student-t-proposal" stream,
    and the log density of each draw: center + L z sqrt(dof / chi2) with
    L L^T = scale (calibration_coverage.py:163-178)."""
    center = np.asarray(center, dtype=float)
    scale = np.asarray(scale, dtype=float)
    dim = center.size
    chol = np.linalg.cholesky(scale)
    precision = np.linalg.inv(scale)
    _, logdet = np.linalg.slogdet(scale)
    log_norm = (math.lgamma((dof + dim) / 2.0) - math.lgamma(dof / 2.0)
                - 0.5 * dim * math.log(dof * math.pi) - 0.5 * logdet)
    samples, log_q = [], {}
    for sid in sample_ids:
        rng = stats.sample_rng(root_seed, sid, "student-t-proposal")
        z = rng.standard_normal(dim)
        chi2 = rng.chisquare(dof)
        offset = chol @ z * math.sqrt(dof / chi2)
        samples.append(Sample(sid, tuple((center + offset).tolist())))
        log_q[sid] = log_norm - 0.5 * (dof + dim) * math.log1p(offset @ precision @ offset / dof)
    return tuple(samples), log_q


@dataclass
class ImportancePosterior:
    solves: int
    weights: np.ndarray
    effective_sample_size: float
    max_weight: float
    theta_mean: np.ndarray
    theta_sd: np.ndarray
    distinct_draws: int
    posterior: EvaluationSet
    predictive: stats.PredictiveSummary
    low: np.ndarray
    high: np.ndarray


# ------------------------------------------------------------ the driver
@dataclass
class CalibrationResult:
    stage1: EvaluationSet
    search_path: list
    optimum: np.ndarray
    final_step: np.ndarray
    grid: EvaluationSet
    laplace: LaplaceFit
    center: np.ndarray
    scale: np.ndarray
    main: ImportancePosterior
    replicate: ImportancePosterior
    coverage: dict = field(default_factory=dict)
    solves: int = 0
    launches: int = 0


class Calibration:
    """The staged calibration over any evaluator that maps a tuple of
    Samples to an EvaluationSet (one GPU batch per call for the engine:
    ``Calibration.for_batch``)."""

    def __init__(self, evaluate: Evaluator, settings: CalibrationSettings | None = None):
        self._evaluate = evaluate
        self.settings = settings or CalibrationSettings()
        self.solves = 0
        self.launches = 0

    @classmethod
    def for_batch(cls, batch: ForwardBatch, settings: CalibrationSettings | None = None,
                  workers: int = 1) -> "Calibration":
        return cls(lambda samples: stats.evaluate_batch(samples, batch, workers=workers), settings)

    def evaluate(self, samples: tuple) -> EvaluationSet:
        """One batch; any failed row stops the calibration."""
        out = self._evaluate(tuple(samples))
        self.solves += len(out.records)
        self.launches += 1
        if out.failure_count:
            raise RuntimeError(f"{out.failure_count} solves failed; the posterior is undefined")
        return out

    # stage 1
    def locate(self) -> tuple:
        s = self.settings
        lower, upper = s.stage1_box
        design = stats.independent_uniform_design(range(s.stage1_draws), lower, upper,
                                                  root_seed=s.seed)
        stage1 = self.evaluate(design)
        return stage1, max(stage1.records, key=lambda r: r.log_likelihood)

    # stage 2
    def pattern_search(self, start, start_loglik: float) -> tuple:
        """Best point of the 8 stencil neighbours replaces the centre when it
        improves on it; otherwise the step halves. Ends after the set
        number of halvings (calibration_coverage.py:121-140)."""
        centre = np.asarray(start, dtype=float)
        best = float(start_loglik)
        step = np.asarray(self.settings.search_step, dtype=float)
        stencil = [(i, j) for i in (-1, 0, 1) for j in (-1, 0, 1) if (i, j) != (0, 0)]
        path = [[*centre.tolist(), best]]
        halvings = 0
        while halvings < self.settings.search_halvings:
            poll = tuple(Sample(k, tuple((centre + step * np.asarray(o)).tolist()))
                         for k, o in enumerate(stencil))
            top = max(self.evaluate(poll).records, key=lambda r: r.log_likelihood)
            if top.log_likelihood > best:
                centre, best = np.asarray(top.theta, dtype=float), top.log_likelihood
                path.append([*centre.tolist(), best])
            else:
                step = step / 2.0
                halvings += 1
        return centre, step, path

    # stage 3
    def laplace(self, optimum, step) -> tuple:
        k = self.settings.grid_half_width
        offsets = np.array([(i, j) for i in range(-k, k + 1) for j in range(-k, k + 1)],
                           dtype=float)
        grid = self.evaluate(tuple(Sample(200 + n, tuple((optimum + step * o).tolist()))
                                   for n, o in enumerate(offsets)))
        scaled = np.array([(np.asarray(r.theta) - optimum) / step for r in grid.records])
        fit = laplace_fit(np.array([r.log_likelihood for r in grid.records]), scaled)
        if np.abs(fit.peak).max() > k:
            raise RuntimeError("the fitted peak lies outside the fitting grid")
        centre = optimum + step * fit.peak
        scale = np.diag(step) @ fit.covariance @ np.diag(step)
        return grid, fit, centre, scale

    # stage 4
    def importance(self, centre, scale, ids: Sequence[int], seed: int,
                   sigma: np.ndarray) -> ImportancePosterior:
        """Weights L/q under a flat prior: the proposal's log density is
        subtracted from each log-likelihood before the prior-design weights
        are formed (calibration_coverage.py:181-225)."""
        s = self.settings
        draws, log_q = student_t_proposal(centre, scale, ids, root_seed=seed, dof=s.proposal_dof)
        evs = self.evaluate(draws)
        ratio = EvaluationSet(tuple(
            dataclasses.replace(r, log_likelihood=r.log_likelihood - log_q[r.sample_id])
            for r in evs.records))
        w = stats.posterior_weights(ratio)
        theta = np.array([r.theta for r in ratio.records])
        mean = w @ theta
        picks = stats.resample_posterior(ratio, range(1000, 1000 + s.posterior_draws),
                                         root_seed=seed + 1)
        # a resampled draw repeats an evaluated coefficient row, whose
        # (deterministic) solve is reused rather than run again
        by_theta = {r.theta: r for r in evs.records}
        posterior = EvaluationSet(tuple(
            stats.Evaluation(p.sample_id, p.theta, "Converged", by_theta[p.theta].predictions,
                             by_theta[p.theta].log_likelihood, by_theta[p.theta].iterations,
                             by_theta[p.theta].backend, "")
            for p in picks))
        predictive = stats.predictive_summary(posterior, root_seed=seed + 2,
                                              observation_sigma=sigma)
        noisy = np.asarray(predictive.samples)
        low, high = np.quantile(noisy, [s.alpha / 2, 1 - s.alpha / 2], axis=0)
        return ImportancePosterior(
            solves=len(evs.records), weights=w, effective_sample_size=float(1.0 / np.sum(w * w)),
            max_weight=float(w.max()), theta_mean=mean,
            theta_sd=np.sqrt(w @ (theta - mean) ** 2), distinct_draws=len({p.theta for p in picks}),
            posterior=posterior, predictive=predictive, low=low, high=high)

    def run(self, truth: np.ndarray, sigma: np.ndarray, group: np.ndarray) -> CalibrationResult:
        """Every stage, then the coverage of the standard and the
        recalibrated 1 - alpha ranges on the held-out points (group 2), with
        group 1 calibrating the log-scale split-conformal correction."""
        s = self.settings
        truth = np.asarray(truth, dtype=float)
        sigma = np.asarray(sigma, dtype=float)
        group = np.asarray(group)
        stage1, first = self.locate()
        optimum, step, path = self.pattern_search(first.theta, first.log_likelihood)
        grid, fit, centre, scale = self.laplace(optimum, step)
        n = s.proposal_draws
        main = self.importance(centre, scale, range(300, 300 + n), s.seed + 1, sigma)
        replicate = self.importance(centre, scale, range(400, 400 + n), s.seed + 11, sigma)
        prediction = np.asarray(main.predictive.latent_mean)
        cal, test = group == 1, group == 2
        log_lo, log_hi = stats.split_conformal_intervals(np.log(truth[cal]),
                                                         np.log(prediction[cal]),
                                                         np.log(prediction), alpha=s.alpha)
        lo, hi = np.exp(log_lo), np.exp(log_hi)
        coverage = {
            "held_out": int(test.sum()),
            "standard_covered": int(np.sum((truth >= main.low) & (truth <= main.high) & test)),
            "recalibrated_covered": int(np.sum((truth >= lo) & (truth <= hi) & test)),
            "replicate_covered": int(np.sum((truth >= replicate.low) & (truth <= replicate.high)
                                            & test)),
            "prediction": prediction, "recalibrated_low": lo, "recalibrated_high": hi,
        }
        return CalibrationResult(stage1, path, optimum, step, grid, fit, centre, scale, main,
                                 replicate, coverage, self.solves, self.launches)


def channel_batch(problem, y: np.ndarray, truth: np.ndarray, likelihood_sigma: np.ndarray,
                  **engine_options) -> ForwardBatch:
    """A ForwardBatch of a channel problem (its mesh, boundary conditions,
    settings and initial state) observed by streamwise velocity at the
    wall distances y with the given observed values and likelihood sigmas;
    points with infinite sigma get sigma 1e6, so their likelihood term is a
    constant (calibration_coverage.py:81-108)."""
    obs = ObservationOperator()
    for yi, ui, si in zip(y, truth, likelihood_sigma):
        obs.add_velocity_profile((0.5, float(yi), 0.0), 0, float(ui),
                                 float(si) if math.isfinite(si) else 1e6)
    p = problem
    return ForwardBatch(p.mesh, InferenceParameterSet.a1_betaStar(), obs,
                        FlowBoundaryConditions.channel_defaults(p.mesh, 1.0, 1.0, 10.0), p.nu,
                        p.settings, Vec3(p.u_init.x, p.u_init.y, 0.0), p.p_init, p.k_init,
                        p.omega_init, **engine_options)
