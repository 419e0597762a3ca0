"""Status-aware statistics over GPU forward-model ensembles: the reference's
statistics layer (rans-uq/python/cfdr/statistics.py) with the same public
functions, arguments, results and errors, evaluated through this engine's
``ForwardBatch`` and ``ForwardModel``.

Failed rows are never admitted: any evaluation that is not Converged (or that
carries an error) makes the posterior undefined, whatever its log-likelihood
sentinel. Every random stream is keyed by (root seed, sample ID, purpose)
through SHA-256, so results do not depend on row order, batching or worker
assignment.
"""
from __future__ import annotations

import hashlib
import math
from collections.abc import Callable, Sequence
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import engine
from .model import ForwardModel

STATUS_NAMES = ("Converged", "Unconverged", "InvalidParameters", "DivergenceDetected", "Unknown")
FAILED_WITHOUT_PREDICTIONS = ("InvalidParameters", "DivergenceDetected", "Unknown")
BACKEND_MODEL = "cfdr_amd.ForwardModel"
BACKEND_BATCH = "cfdr_amd.ForwardBatch"


def _nonnegative_int(value, what: str) -> None:
    ok = isinstance(value, int) and not isinstance(value, bool) and value >= 0
    if not ok:
        raise ValueError(f"{what} must be a nonnegative integer")


def _unique_ids(values: Sequence[int]) -> tuple:
    ids = tuple(values)
    for v in ids:
        _nonnegative_int(v, "sample ID")
    if len(ids) != len(set(ids)):
        raise ValueError("sample IDs must be unique")
    return ids


def sample_rng(root_seed: int, sample_id: int, purpose: str) -> np.random.Generator:
    """The PCG64 stream of (root seed, sample ID, purpose), statistics.py:42-58:
    the SHA-256 digest of "cfdr-statistics-v1:<root>:<id>:<purpose>" read as
    eight little-endian 32-bit words seeds a SeedSequence."""
    _nonnegative_int(root_seed, "root seed")
    _nonnegative_int(sample_id, "sample ID")
    if not purpose:
        raise ValueError("RNG purpose must be nonempty")
    key = f"cfdr-statistics-v1:{root_seed}:{sample_id}:{purpose}".encode("utf-8")
    digest = hashlib.sha256(key).digest()
    words = [int.from_bytes(digest[4 * w:4 * w + 4], "little") for w in range(8)]
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence(words)))


@dataclass(frozen=True)
class Sample:
    sample_id: int
    theta: tuple

    def __post_init__(self):
        _nonnegative_int(self.sample_id, "sample ID")


def independent_uniform_design(sample_ids: Sequence[int], lower, upper, *,
                               root_seed: int) -> tuple:
    """One independent uniform prior draw per ID (statistics.py:69-90)."""
    ids = _unique_ids(sample_ids)
    lo = np.asarray(lower, dtype=float)
    hi = np.asarray(upper, dtype=float)
    valid = (lo.ndim == 1 and lo.size > 0 and hi.shape == lo.shape and
             bool(np.all(np.isfinite(lo))) and bool(np.all(np.isfinite(hi))) and
             bool(np.all(lo < hi)) and bool(np.all(np.isfinite(hi - lo))))
    if not valid:
        raise ValueError("uniform bounds must be finite nonempty vectors with lower < upper")
    _nonnegative_int(root_seed, "root seed")
    out = []
    for i in ids:
        draw = sample_rng(root_seed, i, "prior-design").uniform(lo, hi)
        out.append(Sample(i, tuple(float(v) for v in draw)))
    return tuple(out)


@dataclass(frozen=True)
class Evaluation:
    sample_id: int
    theta: tuple
    status: str
    predictions: tuple
    log_likelihood: float | None
    iterations: int
    backend: str
    error: str

    def __post_init__(self):
        _nonnegative_int(self.sample_id, "sample ID")
        _nonnegative_int(self.iterations, "iteration count")
        if self.status not in STATUS_NAMES:
            raise ValueError("unrecognized evaluation status")
        if not self.backend:
            raise ValueError("evaluation backend must be explicit")

    @property
    def admitted(self) -> bool:
        return self.status == "Converged" and not self.error


@dataclass(frozen=True)
class EvaluationSet:
    records: tuple

    def __post_init__(self):
        _unique_ids([r.sample_id for r in self.records])

    @property
    def failure_count(self) -> int:
        return sum(0 if r.admitted else 1 for r in self.records)


def _record(sample: Sample, result, backend: str, error: str = "") -> Evaluation:
    """One engine row as an Evaluation (statistics.py:134-158): failed rows
    keep no predictions, and a Converged row without finite predictions and
    log-likelihood becomes Unknown."""
    status = str(result.status.name)
    preds = () if status in FAILED_WITHOUT_PREDICTIONS else tuple(float(v) for v in
                                                                  result.predictions)
    loglik = float(result.log_lik) if status == "Converged" else None
    if status == "Converged":
        finite = (len(preds) > 0 and all(math.isfinite(v) for v in preds) and
                  loglik is not None and math.isfinite(loglik))
        if not finite:
            status, loglik, error = "Unknown", None, "nonfinite or missing successful output"
    return Evaluation(sample.sample_id, sample.theta, status, preds, loglik,
                      int(result.simple_iters), backend, error)


def _checked_samples(samples: Sequence[Sample]) -> tuple:
    rows = tuple(samples)
    _unique_ids([s.sample_id for s in rows])
    if rows:
        n = len(rows[0].theta)
        if n == 0 or any(len(s.theta) != n for s in rows):
            raise ValueError("sample parameter vectors must have equal nonzero lengths")
    return rows


def evaluate_scalar(samples: Sequence[Sample], model_factory: Callable[[], Any]) -> EvaluationSet:
    """A fresh ForwardModel with its warm-start cache off for every sample
    (statistics.py:171-184)."""
    out = []
    for s in _checked_samples(samples):
        m = model_factory()
        if not isinstance(m, ForwardModel):
            raise TypeError("model_factory must construct cfdr_amd.model.ForwardModel")
        m.cache().set_byte_budget(0)
        ev = m.evaluate_with_diagnostics(s.theta)
        out.append(_record(s, ev.result, BACKEND_MODEL, ev.error))
    return EvaluationSet(tuple(out))


def evaluate_batch(samples: Sequence[Sample], batch: Any, *, workers: int = 1) -> EvaluationSet:
    """All samples in one ForwardBatch evaluation; failed rows stay aligned
    with their IDs (statistics.py:187-205)."""
    rows = _checked_samples(samples)
    if not isinstance(batch, engine.ForwardBatch):
        raise TypeError("batch must be cfdr_amd.engine.ForwardBatch")
    _nonnegative_int(workers, "worker count")
    if workers < 1:
        raise ValueError("worker count must be positive")
    if not rows:
        return EvaluationSet(())
    res = batch.evaluate(np.asarray([s.theta for s in rows], dtype=float), workers)
    if len(res.rows) != len(rows):
        raise RuntimeError("native batch returned a different row count")
    return EvaluationSet(tuple(_record(s, r.result, BACKEND_BATCH, r.error)
                               for s, r in zip(rows, res.rows)))


def _require_resolved(evaluations: EvaluationSet) -> None:
    if not evaluations.records:
        raise ValueError("at least one evaluation is required")
    n = evaluations.failure_count
    if n:
        raise ValueError(f"{n} unresolved evaluations; posterior is undefined")


def posterior_weights(evaluations: EvaluationSet) -> np.ndarray:
    """Self-normalised likelihood weights of a prior design (statistics.py:217-232)."""
    _require_resolved(evaluations)
    ll = [r.log_likelihood for r in evaluations.records]
    if any(v is None or not math.isfinite(v) for v in ll):
        raise ValueError("successful log likelihoods must be finite")
    top = max(ll)
    w = [math.exp(v - top) for v in ll]
    total = math.fsum(w)
    return np.asarray(w, dtype=float) / total


def resample_posterior(evaluations: EvaluationSet, sample_ids: Sequence[int], *,
                       root_seed: int) -> tuple:
    """Multinomial draws from the importance posterior, one stream per
    output ID over the records in ID order (statistics.py:235-260)."""
    ids = _unique_ids(sample_ids)
    by_id = EvaluationSet(tuple(sorted(evaluations.records, key=lambda r: r.sample_id)))
    w = posterior_weights(by_id)
    _nonnegative_int(root_seed, "root seed")
    n = len(by_id.records)
    return tuple(Sample(i, by_id.records[int(sample_rng(root_seed, i, "posterior-resample")
                                             .choice(n, p=w))].theta) for i in ids)


@dataclass(frozen=True)
class PredictiveSummary:
    sample_ids: tuple
    samples: tuple
    latent_mean: tuple
    latent_variance: tuple
    surrogate_variance: tuple
    discrepancy_variance: tuple
    observation_variance: tuple
    total_variance: tuple
    includes_surrogate_noise: bool
    includes_discrepancy: bool
    includes_observation_noise: bool


def predictive_summary(evaluations: EvaluationSet, *, root_seed: int, observation_sigma=None,
                       surrogate_variance=None, discrepancy_variance=None) -> PredictiveSummary:
    """Equal-weight empirical mixture of the predictions plus independent
    Gaussian surrogate, discrepancy and observation components
    (statistics.py:277-339). The variances are the mixture's exact ones
    (population variance over records in ID order plus the components)."""
    _require_resolved(evaluations)
    _nonnegative_int(root_seed, "root seed")
    latent = np.asarray([r.predictions for r in evaluations.records], dtype=float)
    if latent.ndim != 2 or latent.shape[1] == 0 or not np.all(np.isfinite(latent)):
        raise ValueError("successful predictions must be finite equal-length nonempty vectors")
    dim = latent.shape[1]

    def comp(v):
        a = np.zeros(dim) if v is None else np.asarray(v, dtype=float)
        if a.shape != (dim,) or not np.all(np.isfinite(a)) or np.any(a < 0):
            raise ValueError("noise components must be finite nonnegative output-length vectors")
        return a

    obs_var = comp(observation_sigma) ** 2
    sur_var = comp(surrogate_variance)
    dis_var = comp(discrepancy_variance)
    if not np.all(np.isfinite(obs_var)):
        raise ValueError("observation variance must be finite")
    noisy = latent.copy()
    for row, rec in enumerate(evaluations.records):
        for purpose, var in (("surrogate-noise", sur_var), ("discrepancy", dis_var),
                             ("observation-noise", obs_var)):
            if np.any(var):
                noisy[row] += sample_rng(root_seed, rec.sample_id, purpose).normal(
                    0, np.sqrt(var), dim)
    order = np.argsort([r.sample_id for r in evaluations.records])
    ordered = latent[order]
    mean = ordered.mean(axis=0)
    var = ordered.var(axis=0, ddof=0)
    total = var + sur_var + dis_var + obs_var
    if not np.all(np.isfinite(noisy)) or not np.all(np.isfinite(total)):
        raise ValueError("predictive samples and total variance must be finite")
    return PredictiveSummary(tuple(r.sample_id for r in evaluations.records),
                             tuple(tuple(float(v) for v in row) for row in noisy), tuple(mean),
                             tuple(var), tuple(sur_var), tuple(dis_var), tuple(obs_var),
                             tuple(total), bool(np.any(sur_var)), bool(np.any(dis_var)),
                             bool(np.any(obs_var)))


def _finite_nonempty(values, what: str) -> np.ndarray:
    a = np.asarray(values, dtype=float)
    if a.ndim != 1 or a.size == 0 or not np.all(np.isfinite(a)):
        raise ValueError(f"{what} must be a finite nonempty vector")
    return a


def conformal_quantile(scores, alpha: float) -> float:
    """The ceil((n+1)(1-alpha))-th smallest score, +inf past n (statistics.py:349-355)."""
    s = _finite_nonempty(scores, "conformal scores")
    if not math.isfinite(alpha) or not 0 < alpha < 1 or np.any(s < 0):
        raise ValueError("alpha must lie in (0,1) and scores must be nonnegative")
    k = math.ceil((s.size + 1) * (1 - alpha))
    if k > s.size:
        return math.inf
    return float(np.sort(s)[k - 1])


def split_conformal_intervals(calibration_y, calibration_prediction, test_prediction, *,
                              alpha: float):
    """Absolute-residual split conformal intervals (statistics.py:358-378)."""
    y = _finite_nonempty(calibration_y, "calibration targets")
    f = _finite_nonempty(calibration_prediction, "calibration predictions")
    t = _finite_nonempty(test_prediction, "test predictions")
    if y.shape != f.shape:
        raise ValueError("calibration target and prediction shapes must match")
    r = conformal_quantile(np.abs(y - f), alpha)
    return t - r, t + r


def empirical_crps(samples, truth: float) -> float:
    """CRPS of the equal-weight empirical forecast, M^2 denominator
    (statistics.py:381-396): mean |x - y| minus the sorted-sample form of
    half the mean pairwise distance."""
    x = np.sort(_finite_nonempty(samples, "forecast samples"))
    if not math.isfinite(truth):
        raise ValueError("truth must be finite")
    m = x.size
    c = x - truth
    weights = 2.0 * np.arange(1, m + 1) - m - 1
    value = float(np.mean(np.abs(c)) - np.dot(weights, c) / (m * m))
    if not math.isfinite(value):
        raise ValueError("CRPS arithmetic must remain finite")
    return value


def randomized_ranks(truth, draws, trial_ids: Sequence[int], *, root_seed: int) -> np.ndarray:
    """Per trial and parameter: draws below the truth plus a uniform integer
    in [0, ties] from the (trial ID, parameter) stream (statistics.py:399-426)."""
    t = np.asarray(truth, dtype=float)
    d = np.asarray(draws, dtype=float)
    ids = _unique_ids(trial_ids)
    _nonnegative_int(root_seed, "root seed")
    aligned = (t.ndim == 2 and d.ndim == 3 and t.shape[0] == len(ids) and
               d.shape[0] == len(ids) and d.shape[2] == t.shape[1] and t.size > 0 and
               d.shape[1] > 0 and np.all(np.isfinite(t)) and np.all(np.isfinite(d)))
    if not aligned:
        raise ValueError("truth [trial,parameter] and draws [trial,draw,parameter] must align")
    below = np.sum(d < t[:, None, :], axis=1, dtype=np.int64)
    ties = np.sum(d == t[:, None, :], axis=1)
    for row, tid in enumerate(ids):
        for p in range(t.shape[1]):
            below[row, p] += sample_rng(root_seed, tid, f"rank-tie:{p}").integers(
                int(ties[row, p]) + 1)
    return below


@dataclass(frozen=True)
class RankHistograms:
    counts: np.ndarray
    expected_counts: np.ndarray
    rank_to_bin: np.ndarray


def rank_histograms(ranks, *, n_draws: int, bins: int = 20) -> RankHistograms:
    """Per-parameter histograms of ranks in [0, n_draws] with the exact
    discrete-uniform expectation per bin (statistics.py:435-462)."""
    _nonnegative_int(n_draws, "draw count")
    _nonnegative_int(bins, "bin count")
    r = np.asarray(ranks)
    valid = (n_draws >= 1 and bins >= 1 and r.ndim == 2 and r.size > 0 and
             np.issubdtype(r.dtype, np.integer) and not np.any(r < 0) and
             not np.any(r > n_draws))
    if not valid:
        raise ValueError("ranks must be a nonempty integer [trial,parameter] array in [0,n_draws]")
    nb = min(bins, n_draws + 1)
    to_bin = np.arange(n_draws + 1, dtype=np.int64) * nb // (n_draws + 1)
    per_bin = np.bincount(to_bin, minlength=nb)
    counts = np.stack([np.bincount(to_bin[r[:, j]], minlength=nb)
                       for j in range(r.shape[1])]).astype(np.int64)
    expected = np.tile(per_bin * (r.shape[0] / (n_draws + 1)), (r.shape[1], 1))
    return RankHistograms(counts, expected, to_bin)
