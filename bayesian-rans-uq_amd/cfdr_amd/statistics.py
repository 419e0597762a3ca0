"""Status-aware statistics over ensembles of GPU forward solves.

The public functions, their arguments, results and error texts are those of
the reference's statistics layer (rans-uq/python/cfdr/statistics.py), so a
calibration script written against ``cfdr.statistics`` runs unchanged on
``cfdr_amd.engine.ForwardBatch`` / ``cfdr_amd.model.ForwardModel``.

Three rules shape everything here:

* Admission is by status, never by value. A row counts only when it is
  Converged with finite predictions and log-likelihood and no error text;
  one unresolved row makes any posterior undefined, whatever its
  likelihood sentinel (statistics.py:208-232 in the reference).
* Randomness is addressed, not sequential. Every draw comes from a stream
  named by (root seed, sample ID, purpose), so a result does not depend on
  row order, batching or which worker ran what.
* Variances are exact properties of the stated mixture, not sample
  estimates of it.
"""
from __future__ import annotations

import hashlib
import math
from collections.abc import Callable, Sequence
from dataclasses import dataclass
from typing import Any, NamedTuple

import numpy as np

from . import engine
from .model import ForwardModel

BACKEND_MODEL = "cfdr_amd.ForwardModel"
BACKEND_BATCH = "cfdr_amd.ForwardBatch"

# Evaluation statuses, the reference's names (bindings.cpp:464-469)
STATUS_NAMES = ("Converged", "Unconverged", "InvalidParameters", "DivergenceDetected", "Unknown")
# statuses whose rows carry no predictions at all (their NaN placeholders are dropped)
FAILED_WITHOUT_PREDICTIONS = ("InvalidParameters", "DivergenceDetected", "Unknown")
_NONFINITE_SUCCESS = "nonfinite or missing successful output"


# ------------------------------------------------------------------ checks
def _count(value, what: str) -> int:
    """A nonnegative Python integer (bool excluded) or ValueError."""
    if type(value) is bool or not isinstance(value, int) or value < 0:
        raise ValueError(f"{what} must be a nonnegative integer")
    return value


def _distinct_ids(values: Sequence[int]) -> tuple:
    ids = tuple(_count(v, "sample ID") for v in values)
    if len(frozenset(ids)) < len(ids):
        raise ValueError("sample IDs must be unique")
    return ids


def _finite_vector(values, what: str) -> np.ndarray:
    a = np.asarray(values, dtype=float)
    if a.ndim == 1 and a.size and np.isfinite(a).all():
        return a
    raise ValueError(f"{what} must be a finite nonempty vector")


# ----------------------------------------------------------------- streams
def _stream(root_seed: int, sample_id: int, purpose: str) -> np.random.Generator:
    """The stream construction every consumer of these results relies on:
    SHA-256 of "cfdr-statistics-v1:<root>:<id>:<purpose>" read as eight
    little-endian 32-bit words seeds a SeedSequence for PCG64
    (statistics.py:42-58). Callers have validated the fields."""
    digest = hashlib.sha256(
        f"cfdr-statistics-v1:{root_seed}:{sample_id}:{purpose}".encode("utf-8")).digest()
    words = np.frombuffer(digest, dtype="<u4").tolist()
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence(words)))


def sample_rng(root_seed: int, sample_id: int, purpose: str) -> np.random.Generator:
    """The PCG64 stream of (root seed, sample ID, purpose)."""
    _count(root_seed, "root seed")
    _count(sample_id, "sample ID")
    if not purpose:
        raise ValueError("RNG purpose must be nonempty")
    return _stream(root_seed, sample_id, purpose)


# ----------------------------------------------------------------- designs
@dataclass(frozen=True)
class Sample:
    sample_id: int
    theta: tuple

    def __post_init__(self):
        _count(self.sample_id, "sample ID")


def independent_uniform_design(sample_ids: Sequence[int], lower, upper, *,
                               root_seed: int) -> tuple:
    """One independent uniform draw in [lower, upper) per ID, each from its
    own "prior-design" stream (not a Latin hypercube)."""
    ids = _distinct_ids(sample_ids)
    lo = np.asarray(lower, dtype=float)
    hi = np.asarray(upper, dtype=float)
    with np.errstate(over="ignore", invalid="ignore"):  # an infinite width is refused below
        width = hi - lo if hi.shape == lo.shape else None
    if not (lo.ndim == 1 and lo.size and width is not None and np.isfinite(lo).all()
            and np.isfinite(hi).all() and np.isfinite(width).all() and (width > 0).all()):
        raise ValueError("uniform bounds must be finite nonempty vectors with lower < upper")
    _count(root_seed, "root seed")
    return tuple(Sample(i, tuple(_stream(root_seed, i, "prior-design").uniform(lo, hi).tolist()))
                 for i in ids)


# ------------------------------------------------------------- evaluations
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
        _count(self.sample_id, "sample ID")
        _count(self.iterations, "iteration count")
        if self.status not in STATUS_NAMES:
            raise ValueError("unrecognized evaluation status")
        if not self.backend:
            raise ValueError("evaluation backend must be explicit")

    @property
    def admitted(self) -> bool:
        """Converged and without error text: the only rows a posterior uses."""
        return not self.error and self.status == "Converged"


@dataclass(frozen=True)
class EvaluationSet:
    records: tuple

    def __post_init__(self):
        _distinct_ids([r.sample_id for r in self.records])

    @property
    def failure_count(self) -> int:
        return len(self.records) - sum(r.admitted for r in self.records)

    def _canonical(self) -> tuple:
        """The records in sample-ID order (the order posterior quantities
        are formed in, so they do not depend on how rows arrived)."""
        return tuple(sorted(self.records, key=lambda r: r.sample_id))


def _from_engine_row(sample: Sample, row, backend: str, error: str = "") -> Evaluation:
    """One engine row as an Evaluation. Rows of the failed statuses keep
    their status and error but no predictions; a Converged row whose
    predictions or log-likelihood are missing or nonfinite is recorded as
    Unknown, so it can never be admitted (statistics.py:134-158)."""
    status = str(row.status.name)
    iterations = int(row.simple_iters)
    if status in FAILED_WITHOUT_PREDICTIONS:
        return Evaluation(sample.sample_id, sample.theta, status, (), None, iterations,
                          backend, error)
    predictions = tuple(float(v) for v in row.predictions)
    if status != "Converged":
        return Evaluation(sample.sample_id, sample.theta, status, predictions, None,
                          iterations, backend, error)
    loglik = float(row.log_lik)
    if predictions and math.isfinite(loglik) and all(map(math.isfinite, predictions)):
        return Evaluation(sample.sample_id, sample.theta, status, predictions, loglik,
                          iterations, backend, error)
    return Evaluation(sample.sample_id, sample.theta, "Unknown", predictions, None, iterations,
                      backend, _NONFINITE_SUCCESS)


def _design_rows(samples: Sequence[Sample]) -> tuple:
    rows = tuple(samples)
    _distinct_ids([s.sample_id for s in rows])
    widths = {len(s.theta) for s in rows}
    if rows and (len(widths) != 1 or 0 in widths):
        raise ValueError("sample parameter vectors must have equal nonzero lengths")
    return rows


def evaluate_scalar(samples: Sequence[Sample], model_factory: Callable[[], Any]) -> EvaluationSet:
    """Each sample on a fresh ForwardModel whose warm-start cache is off
    (statistics.py:171-184)."""
    out = []
    for s in _design_rows(samples):
        model = model_factory()
        if not isinstance(model, ForwardModel):
            raise TypeError("model_factory must construct cfdr_amd.model.ForwardModel")
        model.cache().set_byte_budget(0)
        diag = model.evaluate_with_diagnostics(s.theta)
        out.append(_from_engine_row(s, diag.result, BACKEND_MODEL, diag.error))
    return EvaluationSet(tuple(out))


def evaluate_batch(samples: Sequence[Sample], batch: Any, *, workers: int = 1) -> EvaluationSet:
    """Every sample in one ForwardBatch evaluation (one GPU launch); failed
    rows stay aligned with their IDs (statistics.py:187-205)."""
    rows = _design_rows(samples)
    if not isinstance(batch, engine.ForwardBatch):
        raise TypeError("batch must be cfdr_amd.engine.ForwardBatch")
    if _count(workers, "worker count") == 0:
        raise ValueError("worker count must be positive")
    if not rows:
        return EvaluationSet(())
    native = batch.evaluate(np.asarray([s.theta for s in rows], dtype=float), workers).rows
    if len(native) != len(rows):
        raise RuntimeError("native batch returned a different row count")
    return EvaluationSet(tuple(_from_engine_row(s, n.result, BACKEND_BATCH, n.error)
                               for s, n in zip(rows, native)))


# ------------------------------------------------------ importance posterior
def _resolved(evaluations: EvaluationSet) -> tuple:
    """The records, once every one of them is admitted."""
    records = evaluations.records
    if not records:
        raise ValueError("at least one evaluation is required")
    unresolved = evaluations.failure_count
    if unresolved:
        raise ValueError(f"{unresolved} unresolved evaluations; posterior is undefined")
    return records


def posterior_weights(evaluations: EvaluationSet) -> np.ndarray:
    """Self-normalised likelihood weights of a design drawn from the prior:
    w_i = exp(l_i - max l) / sum_j exp(l_j - max l), in record order."""
    loglik = [r.log_likelihood for r in _resolved(evaluations)]
    if not all(v is not None and math.isfinite(v) for v in loglik):
        raise ValueError("successful log likelihoods must be finite")
    peak = max(loglik)
    unnormalised = [math.exp(v - peak) for v in loglik]
    return np.asarray(unnormalised, dtype=float) / math.fsum(unnormalised)


def resample_posterior(evaluations: EvaluationSet, sample_ids: Sequence[int], *,
                       root_seed: int) -> tuple:
    """Independent multinomial draws from the importance posterior: output
    ID k picks a record of the ID-ordered set with its own
    "posterior-resample" stream, so the draws do not depend on input order,
    output order or chunking."""
    out_ids = _distinct_ids(sample_ids)
    canonical = EvaluationSet(evaluations._canonical())
    weights = posterior_weights(canonical)
    _count(root_seed, "root seed")
    pool = canonical.records
    picks = [int(_stream(root_seed, k, "posterior-resample").choice(len(pool), p=weights))
             for k in out_ids]
    return tuple(Sample(k, pool[j].theta) for k, j in zip(out_ids, picks))


# --------------------------------------------------- predictive distribution
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


class _Noise(NamedTuple):
    purpose: str          # stream purpose of this component's draws
    variance: np.ndarray  # per output


def _noise_variance(value, n_out: int) -> np.ndarray:
    if value is None:
        return np.zeros(n_out)
    v = np.asarray(value, dtype=float)
    if v.shape == (n_out,) and np.isfinite(v).all() and (v >= 0).all():
        return v
    raise ValueError("noise components must be finite nonnegative output-length vectors")


def predictive_summary(evaluations: EvaluationSet, *, root_seed: int, observation_sigma=None,
                       surrogate_variance=None, discrepancy_variance=None) -> PredictiveSummary:
    """The equal-weight mixture of the admitted predictions plus independent
    Gaussian surrogate, discrepancy and observation components
    (statistics.py:277-339). samples holds one noisy draw per record (record
    order; each component from that record's own stream); the variances are
    the mixture's exact ones: the population variance of the predictions
    over records in ID order plus each component's variance."""
    records = _resolved(evaluations)
    _count(root_seed, "root seed")
    latent = np.asarray([r.predictions for r in records], dtype=float)
    if not (latent.ndim == 2 and latent.shape[1] and np.isfinite(latent).all()):
        raise ValueError("successful predictions must be finite equal-length nonempty vectors")
    n_out = latent.shape[1]
    sigma = _noise_variance(observation_sigma, n_out)
    with np.errstate(over="ignore"):  # an infinite variance is refused below
        sigma_sq = sigma * sigma
    components = (_Noise("surrogate-noise", _noise_variance(surrogate_variance, n_out)),
                  _Noise("discrepancy", _noise_variance(discrepancy_variance, n_out)),
                  _Noise("observation-noise", sigma_sq))
    if not np.isfinite(components[2].variance).all():
        raise ValueError("observation variance must be finite")
    active = [c for c in components if c.variance.any()]
    draws = latent.copy()
    for row, rec in zip(draws, records):
        for c in active:
            row += _stream(root_seed, rec.sample_id, c.purpose).normal(
                0.0, np.sqrt(c.variance), n_out)
    by_id = latent[np.argsort([r.sample_id for r in records], kind="stable")]
    with np.errstate(over="ignore", invalid="ignore"):  # overflow is reported below
        mean, spread = by_id.mean(axis=0), by_id.var(axis=0)
    # summed in the component order (surrogate, discrepancy, observation)
    total = spread + components[0].variance + components[1].variance + components[2].variance
    if not (np.isfinite(draws).all() and np.isfinite(total).all()):
        raise ValueError("predictive samples and total variance must be finite")
    sur, dis, obs = (tuple(c.variance.tolist()) for c in components)
    return PredictiveSummary(
        sample_ids=tuple(r.sample_id for r in records),
        samples=tuple(tuple(row.tolist()) for row in draws),
        latent_mean=tuple(mean.tolist()), latent_variance=tuple(spread.tolist()),
        surrogate_variance=sur, discrepancy_variance=dis, observation_variance=obs,
        total_variance=tuple(total.tolist()),
        includes_surrogate_noise=bool(components[0].variance.any()),
        includes_discrepancy=bool(components[1].variance.any()),
        includes_observation_noise=bool(components[2].variance.any()))


# ------------------------------------------------ conformal, scores, ranks
def conformal_quantile(scores, alpha: float) -> float:
    """The k-th smallest score with k = ceil((n + 1)(1 - alpha)), or +inf
    when k exceeds n (statistics.py:349-355)."""
    s = _finite_vector(scores, "conformal scores")
    if not (math.isfinite(alpha) and 0.0 < alpha < 1.0) or (s < 0).any():
        raise ValueError("alpha must lie in (0,1) and scores must be nonnegative")
    k = math.ceil((s.size + 1) * (1.0 - alpha))
    return math.inf if k > s.size else float(np.partition(s, k - 1)[k - 1])


def split_conformal_intervals(calibration_y, calibration_prediction, test_prediction, *,
                              alpha: float):
    """Absolute-residual split conformal intervals: test predictions plus or
    minus the conformal quantile of the calibration residuals
    (statistics.py:358-378)."""
    y = _finite_vector(calibration_y, "calibration targets")
    fit = _finite_vector(calibration_prediction, "calibration predictions")
    test = _finite_vector(test_prediction, "test predictions")
    if y.shape != fit.shape:
        raise ValueError("calibration target and prediction shapes must match")
    radius = conformal_quantile(np.abs(y - fit), alpha)
    return test - radius, test + radius


def empirical_crps(samples, truth: float) -> float:
    """CRPS of the equal-weight M-member forecast with the M^2 denominator,
    E|X - y| - E|X - X'| / 2 (statistics.py:381-396). The pair term uses
    the sorted members: sum over pairs i < j of (x_j - x_i) is
    sum_j ((j - 1) x_j - (x_1 + ... + x_{j-1})), a prefix-sum pass instead
    of the M^2 pairwise array; members are shifted by the truth first, which
    changes no difference and keeps large offsets from cancelling."""
    x = np.sort(_finite_vector(samples, "forecast samples"))
    if not math.isfinite(truth):
        raise ValueError("truth must be finite")
    m = x.size
    with np.errstate(over="ignore", invalid="ignore"):  # overflow is reported below
        c = x - truth
        before = np.concatenate(([0.0], np.cumsum(c)[:-1]))
        pair_sum = float(np.sum(np.arange(m) * c - before))
        value = float(np.mean(np.abs(c))) - pair_sum / (m * m)
    if not math.isfinite(value):
        raise ValueError("CRPS arithmetic must remain finite")
    return value


def randomized_ranks(truth, draws, trial_ids: Sequence[int], *, root_seed: int) -> np.ndarray:
    """Rank of each trial's truth among its draws, per parameter: the draws
    strictly below it plus a uniform integer in [0, ties] from the stream of
    (trial ID, "rank-tie:<parameter>") (statistics.py:399-426)."""
    t = np.asarray(truth, dtype=float)
    d = np.asarray(draws, dtype=float)
    ids = _distinct_ids(trial_ids)
    _count(root_seed, "root seed")
    n = len(ids)
    if not (t.ndim == 2 and d.ndim == 3 and t.shape[0] == n == d.shape[0] and
            d.shape[2] == t.shape[1] and t.size and d.shape[1] and np.isfinite(t).all() and
            np.isfinite(d).all()):
        raise ValueError("truth [trial,parameter] and draws [trial,draw,parameter] must align")
    below = (d < t[:, None, :]).sum(axis=1, dtype=np.int64)
    ties = (d == t[:, None, :]).sum(axis=1)
    for (trial, param), k in np.ndenumerate(ties):
        below[trial, param] += _stream(root_seed, ids[trial], f"rank-tie:{param}").integers(
            int(k) + 1)
    return below


@dataclass(frozen=True)
class RankHistograms:
    counts: np.ndarray
    expected_counts: np.ndarray
    rank_to_bin: np.ndarray


def rank_histograms(ranks, *, n_draws: int, bins: int = 20) -> RankHistograms:
    """Per-parameter histograms of ranks in [0, n_draws]: rank r falls in bin
    floor(r * B / (n_draws + 1)) with B = min(bins, n_draws + 1), and each
    bin's expected count under uniform ranks is its number of ranks times
    trials / (n_draws + 1) (statistics.py:435-462). No p-value is given."""
    _count(n_draws, "draw count")
    _count(bins, "bin count")
    r = np.asarray(ranks)
    if not (n_draws >= 1 and bins >= 1 and r.ndim == 2 and r.size and
            np.issubdtype(r.dtype, np.integer) and r.min() >= 0 and r.max() <= n_draws):
        raise ValueError("ranks must be a nonempty integer [trial,parameter] array in [0,n_draws]")
    n_bins = min(bins, n_draws + 1)
    to_bin = (np.arange(n_draws + 1, dtype=np.int64) * n_bins) // (n_draws + 1)
    counts = np.zeros((r.shape[1], n_bins), dtype=np.int64)
    for param in range(r.shape[1]):
        np.add.at(counts[param], to_bin[r[:, param]], 1)
    per_bin = np.bincount(to_bin, minlength=n_bins)
    expected = np.broadcast_to(per_bin * (r.shape[0] / (n_draws + 1)),
                               (r.shape[1], n_bins)).copy()
    return RankHistograms(counts, expected, to_bin)
