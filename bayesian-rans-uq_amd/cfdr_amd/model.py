"""Single-row forward model with a device-resident warm-start cache
(include/rans_amd.h, rq_model_*).

``ForwardModel`` mirrors the reference's ForwardModel and its Python binding
(rans-uq/include/ForwardModel.hpp:191-247, rans-uq/src/ForwardModel.cpp:35-137,
rans-uq/python/bindings.cpp:808-828,917-938): ``evaluate(theta)`` and
``evaluate_with_diagnostics(theta)`` solve one coefficient row, starting from
the nearest cached converged state when one lies within the cache threshold
(``used_warm_start``), and store the state of every converged row. The cached
states stay in device memory. ``evaluate_batch`` is this engine's addition:
several rows look up the cache at once and run in one launch.

``WarmStartCache`` is the cache's view (size, clear, byte budget, retained
bytes); ``HostWarmCache`` is the same bookkeeping without a GPU, entries
tagged by integers.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import engine
from .config import (FlowBoundaryConditions, InferenceParameterSet, ObservationOperator,
                     SolverSettings, Vec3)
from .mesh import Mesh

DEFAULT_MAX_SIZE = 50
DEFAULT_THRESHOLD = 0.5
DEFAULT_BYTE_BUDGET = 64 * 1024 * 1024


class RqCacheOptions(C.Structure):
    _fields_ = [("max_size", C.c_int32), ("threshold", C.c_double),
                ("byte_budget", C.c_int64)]


def _bind(lib):
    if getattr(lib, "_model_bound", False):
        return lib
    ep = [C.POINTER(engine.RqMesh), C.POINTER(engine.RqParams), C.POINTER(engine.RqObs),
          C.POINTER(engine.RqBc), C.c_double, C.POINTER(engine.RqSettings), engine._dp]
    lib.rq_model_create.argtypes = ep + [C.POINTER(RqCacheOptions), C.c_char_p, C.c_size_t]
    lib.rq_model_create.restype = C.c_void_p
    for name in ("rq_model_evaluate", "rq_model_evaluate_batch"):
        getattr(lib, name).argtypes = [C.c_void_p, engine._dp, C.c_int64, C.c_int32,
                                       C.POINTER(engine.RqBatchOut), engine._u8p, C.c_char_p,
                                       C.c_size_t]
        getattr(lib, name).restype = C.c_int
    lib.rq_model_engine.argtypes = [C.c_void_p]
    lib.rq_model_engine.restype = C.c_void_p
    lib.rq_model_cache_size.argtypes = [C.c_void_p]
    lib.rq_model_cache_size.restype = C.c_int32
    for name in ("rq_model_cache_retained_bytes", "rq_model_cache_byte_budget"):
        getattr(lib, name).argtypes = [C.c_void_p]
        getattr(lib, name).restype = C.c_int64
    lib.rq_model_cache_set_byte_budget.argtypes = [C.c_void_p, C.c_int64]
    lib.rq_model_cache_set_byte_budget.restype = C.c_int
    lib.rq_model_cache_clear.argtypes = [C.c_void_p]
    lib.rq_model_cache_clear.restype = C.c_int
    lib.rq_model_destroy.argtypes = [C.c_void_p]
    lib.rq_model_destroy.restype = None
    lib.rq_warm_cache_create.argtypes = [C.c_int32, C.c_double, C.c_int64, C.c_char_p,
                                         C.c_size_t]
    lib.rq_warm_cache_create.restype = C.c_void_p
    lib.rq_warm_cache_store.argtypes = [C.c_void_p, engine._dp, C.c_int32, C.c_int64, C.c_int64]
    lib.rq_warm_cache_store.restype = C.c_int32
    lib.rq_warm_cache_nearest.argtypes = [C.c_void_p, engine._dp, C.c_int32]
    lib.rq_warm_cache_nearest.restype = C.c_int64
    lib.rq_warm_cache_size.argtypes = [C.c_void_p]
    lib.rq_warm_cache_size.restype = C.c_int32
    lib.rq_warm_cache_retained_bytes.argtypes = [C.c_void_p]
    lib.rq_warm_cache_retained_bytes.restype = C.c_int64
    lib.rq_warm_cache_set_byte_budget.argtypes = [C.c_void_p, C.c_int64]
    lib.rq_warm_cache_set_byte_budget.restype = C.c_int
    lib.rq_warm_cache_clear.argtypes = [C.c_void_p]
    lib.rq_warm_cache_clear.restype = C.c_int
    lib.rq_warm_cache_entry_bytes.argtypes = [C.c_int32, C.c_int32, C.c_int32]
    lib.rq_warm_cache_entry_bytes.restype = C.c_int64
    lib.rq_warm_cache_destroy.argtypes = [C.c_void_p]
    lib.rq_warm_cache_destroy.restype = None
    lib._model_bound = True
    return lib


def _row(theta) -> np.ndarray:
    return np.ascontiguousarray(theta, dtype=np.float64).ravel()


@dataclass
class ForwardEvaluation:
    """ForwardEvaluationResult (ForwardModel.hpp:188-195) without the
    convergence history object: its scalars are fields here."""

    result: engine.EvaluationResult
    used_warm_start: bool
    final_worst_measure: float
    contraction_rate_u: float
    stalled: bool
    error: str
    krylov_iterations: tuple


class WarmStartCache:
    """WarmStartCache's Python surface (bindings.cpp:823-828)."""

    def __init__(self, model: "ForwardModel"):
        self._m = model

    def size(self) -> int:
        return int(self._m.lib.rq_model_cache_size(self._m._h))

    def clear(self) -> None:
        self._m.lib.rq_model_cache_clear(self._m._h)

    def set_byte_budget(self, nbytes: int) -> None:
        if nbytes < 0:
            raise ValueError("byte budget must be nonnegative")
        self._m.lib.rq_model_cache_set_byte_budget(self._m._h, int(nbytes))

    def byte_budget(self) -> int:
        return int(self._m.lib.rq_model_cache_byte_budget(self._m._h))

    def retained_bytes(self) -> int:
        return int(self._m.lib.rq_model_cache_retained_bytes(self._m._h))


class ForwardModel:
    """ForwardModel on one GPU with a device-resident warm-start cache."""

    def __init__(self, mesh: Mesh, param_set: InferenceParameterSet,
                 obs_op: ObservationOperator, bcs: FlowBoundaryConditions, nu: float,
                 settings: SolverSettings = None, u_init: Vec3 = Vec3(1.0, 0.0, 0.0),
                 p_init: float = 0.0, k_init: float = 1e-4, omega_init: float = 1.0,
                 cache_max_size: int = DEFAULT_MAX_SIZE,
                 cache_threshold: float = DEFAULT_THRESHOLD,
                 cache_byte_budget: int = DEFAULT_BYTE_BUDGET, block_size: int = 0,
                 grid_blocks: int = 0):
        self.lib = _bind(engine.load_library())
        pr = engine.Problem(mesh, param_set, obs_op, bcs, nu, settings, u_init, p_init, k_init,
                            omega_init, False, block_size, False, grid_blocks, "ForwardModel")
        self._param_set = param_set
        self.n_obs = obs_op.n_obs()
        self.n_cells, self.n_faces = mesh.n_cells, int(pr.m.n_faces)
        opt = RqCacheOptions(cache_max_size, cache_threshold, cache_byte_budget)
        err = C.create_string_buffer(512)
        h = self.lib.rq_model_create(*pr.args(), C.byref(opt), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h

    def close(self) -> None:
        """Releases the native object now (rq_model_destroy) instead of at garbage
        collection."""
        h = getattr(self, "_h", None)
        if h and getattr(self, "lib", None) is not None:
            self._h = None
            self.lib.rq_model_destroy(h)

    def __del__(self):
        self.close()

    def param_set(self) -> InferenceParameterSet:
        return self._param_set

    def cache(self) -> WarmStartCache:
        return WarmStartCache(self)

    def _run(self, theta, batch: bool):
        theta = np.ascontiguousarray(np.atleast_2d(np.asarray(theta, dtype=np.float64)))
        n_params = self._param_set.n_active()
        if theta.shape[1] != n_params:
            raise ValueError("theta must have n_parameters entries")
        n = theta.shape[0]
        bb = engine.BatchBuffers(n, self.n_obs, self.n_cells, self.n_faces)
        used = np.zeros(n, np.uint8)
        err = C.create_string_buffer(512)
        fn = self.lib.rq_model_evaluate_batch if batch else self.lib.rq_model_evaluate
        rc = fn(self._h, theta.ctypes.data_as(engine._dp), n, n_params, C.byref(bb.out),
                used.ctypes.data_as(engine._u8p), err, 512)
        engine.BatchBuffers.check(rc, err)
        return bb.result(), used.astype(bool)

    def evaluate_with_diagnostics(self, theta) -> ForwardEvaluation:
        res, used = self._run(_row(theta)[None, :], False)
        r = res.rows[0]
        return ForwardEvaluation(r.result, bool(used[0]), r.final_worst_measure,
                                 r.contraction_rate_u, r.stalled, r.error,
                                 tuple(int(v) for v in res.krylov_iterations[0]))

    def evaluate(self, theta) -> engine.EvaluationResult:
        return self.evaluate_with_diagnostics(theta).result

    def penalized_log_likelihood(self, theta) -> float:
        """ForwardModel::penalizedLogLikelihood (ForwardModel.cpp:140-143)."""
        r = self.evaluate(theta)
        return r.log_lik if r.status == engine.EvaluationStatus.Converged else -1e30

    def evaluate_sequence(self, theta):
        """Rows one after another, each seeing the states the earlier rows
        stored; returns (ForwardBatchResult, used_warm_start mask)."""
        return self._run(theta, False)

    def evaluate_batch(self, theta):
        """Rows in one launch, each warm-started from the cache as it stands
        at the call; converged rows are stored afterwards in row order.
        Returns (ForwardBatchResult, used_warm_start mask)."""
        return self._run(theta, True)


class HostWarmCache:
    """The warm-start cache's bookkeeping without a GPU (rq_warm_cache_*):
    entries carry an integer tag in place of a flow state."""

    def __init__(self, max_size: int = DEFAULT_MAX_SIZE, threshold: float = DEFAULT_THRESHOLD,
                 byte_budget: int = DEFAULT_BYTE_BUDGET):
        self.lib = _bind(engine.load_library())
        err = C.create_string_buffer(512)
        h = self.lib.rq_warm_cache_create(int(max_size), float(threshold), int(byte_budget), err,
                                          512)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h

    def close(self) -> None:
        """Releases the native object now (rq_warm_cache_destroy) instead of at garbage
        collection."""
        h = getattr(self, "_h", None)
        if h and getattr(self, "lib", None) is not None:
            self._h = None
            self.lib.rq_warm_cache_destroy(h)

    def __del__(self):
        self.close()

    @staticmethod
    def entry_bytes(n_theta: int, n_cells: int, n_boundary: int) -> int:
        lib = _bind(engine.load_library())
        return int(lib.rq_warm_cache_entry_bytes(n_theta, n_cells, n_boundary))

    def store(self, theta, tag: int, nbytes: int) -> bool:
        t = _row(theta)
        return self.lib.rq_warm_cache_store(self._h, t.ctypes.data_as(engine._dp), t.size,
                                            int(tag), int(nbytes)) == 1

    def nearest(self, theta):
        t = _row(theta)
        tag = self.lib.rq_warm_cache_nearest(self._h, t.ctypes.data_as(engine._dp), t.size)
        return None if tag < 0 else int(tag)

    def size(self) -> int:
        return int(self.lib.rq_warm_cache_size(self._h))

    def retained_bytes(self) -> int:
        return int(self.lib.rq_warm_cache_retained_bytes(self._h))

    def set_byte_budget(self, nbytes: int) -> None:
        self.lib.rq_warm_cache_set_byte_budget(self._h, int(nbytes))

    def clear(self) -> None:
        self.lib.rq_warm_cache_clear(self._h)
