"""Continuous-batching front end of the engine (include/rans_amd.h, rq_server_*).

A ``Server`` holds registered problems (each one the inputs of a reference
``ForwardBatch``: mesh, parameter set, observations, boundary conditions,
viscosity, settings, initial state) and accepts single-row requests at any
time. A persistent grid on the GPU claims requests as its slots free up, so
requests that arrive while others run are admitted and overlap on the device.
Each finished request carries the fields of a reference ``ForwardBatchRow``
(rans-uq/include/ForwardBatch.hpp:8-14) with the same status and
log-likelihood conventions.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import engine
from .config import (FlowBoundaryConditions, InferenceParameterSet, ObservationOperator,
                     SolverSettings, Vec3)
from .mesh import Mesh


class RqServeOptions(C.Structure):
    _fields_ = [("max_slots", C.c_int32), ("hbm_budget_bytes", C.c_int64),
                ("ring_size", C.c_int32), ("idle_exit_seconds", C.c_double)]


class RqServeRow(C.Structure):
    _fields_ = [("status", C.c_int32), ("loglik", C.c_double),
                ("predictions", C.POINTER(C.c_double)), ("iterations", C.c_int32),
                ("final_worst_measure", C.c_double), ("contraction_rate_u", C.c_double),
                ("stalled", C.c_uint8), ("converged", C.c_uint8),
                ("krylov_iterations", C.c_int64 * 5), ("error", C.c_char * 256)]


SERVER_EXPORTS = ["rq_server_create", "rq_server_add_problem", "rq_server_submit",
                  "rq_server_poll", "rq_server_stats", "rq_server_destroy"]


def _bind(lib):
    if getattr(lib, "_serve_bound", False):
        return lib
    lib.rq_server_create.argtypes = [C.POINTER(RqServeOptions), C.c_char_p, C.c_size_t]
    lib.rq_server_create.restype = C.c_void_p
    lib.rq_server_add_problem.argtypes = [C.c_void_p, C.POINTER(engine.RqMesh),
                                          C.POINTER(engine.RqParams), C.POINTER(engine.RqObs),
                                          C.POINTER(engine.RqBc), C.c_double,
                                          C.POINTER(engine.RqSettings), engine._dp, C.c_char_p,
                                          C.c_size_t]
    lib.rq_server_add_problem.restype = C.c_int32
    lib.rq_server_submit.argtypes = [C.c_void_p, C.c_int32, engine._dp, C.c_int32, C.c_char_p,
                                     C.c_size_t]
    lib.rq_server_submit.restype = C.c_int64
    lib.rq_server_poll.argtypes = [C.c_void_p, C.c_int64, C.POINTER(RqServeRow),
                                   C.POINTER(C.c_int32), C.c_char_p, C.c_size_t]
    lib.rq_server_poll.restype = C.c_int
    lib.rq_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.rq_server_stats.restype = C.c_int
    lib.rq_server_destroy.argtypes = [C.c_void_p]
    lib.rq_server_destroy.restype = None
    lib._serve_bound = True
    return lib


@dataclass
class ServedRow:
    """One finished request: a ForwardBatchRow plus its Krylov counts."""

    status: engine.EvaluationStatus
    log_lik: float
    predictions: np.ndarray
    simple_iters: int
    final_worst_measure: float
    contraction_rate_u: float
    stalled: bool
    error: str
    krylov_iterations: tuple


class Server:
    """Registered problems plus a device-resident request loop.

    ``max_slots`` bounds the concurrent solves (the device's CU count bounds
    it too); ``hbm_budget_bytes`` bounds the arena memory of the slots, so
    the slot count is min(max_slots, budget // arena bytes per slot, CUs).
    """

    def __init__(self, max_slots: int = 256, hbm_budget_bytes: int = 0, ring_size: int = 4096,
                 idle_exit_seconds: float = 2.0):
        self.lib = _bind(engine.load_library())
        opt = RqServeOptions(max_slots, hbm_budget_bytes, ring_size, idle_exit_seconds)
        err = C.create_string_buffer(512)
        h = self.lib.rq_server_create(C.byref(opt), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h
        self._nobs = []
        self._nparams = []

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self.lib.rq_server_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_problem(self, mesh: Mesh, param_set: InferenceParameterSet,
                    obs_op: ObservationOperator, bcs: FlowBoundaryConditions, nu: float,
                    settings: SolverSettings = None, u_init: Vec3 = Vec3(1.0, 0.0, 0.0),
                    p_init: float = 0.0, k_init: float = 1e-4, omega_init: float = 1.0) -> int:
        settings = settings if settings is not None else SolverSettings()
        settings.validate()
        param_set.validate()
        bcs.validate(mesh)
        if not mesh.wall_distance_valid():
            raise ValueError("add_problem needs compute_wall_distance before registration")
        mesh, bcs, folded = engine.planar_problem(mesh, bcs)
        keep = engine._Keep()
        m = engine.mesh_struct(mesh, keep, folded)
        b = engine.RqBc()
        ut, uv, pt, pv, kt, kv, wt, wv = bcs.arrays()
        b.u_type, b.p_type, b.k_type, b.w_type = keep.i(ut), keep.i(pt), keep.i(kt), keep.i(wt)
        b.u_val, b.p_val, b.k_val, b.w_val = keep.d(uv), keep.d(pv), keep.d(kv), keep.d(wv)
        p = engine.RqParams()
        p.n_active = param_set.n_active()
        p.active_indices = keep.i(param_set.active_indices)
        p.base = keep.d(param_set.base)
        p.lower, p.upper = keep.d(param_set.lower_bounds()), keep.d(param_set.upper_bounds())
        o = engine.obs_struct(obs_op.prepare(mesh), keep)
        s = engine.settings_struct(settings)
        init = np.asarray([u_init.x, u_init.y, p_init, k_init, omega_init], float)
        err = C.create_string_buffer(512)
        pid = self.lib.rq_server_add_problem(self._h, C.byref(m), C.byref(p), C.byref(o),
                                             C.byref(b), float(nu), C.byref(s),
                                             init.ctypes.data_as(engine._dp), err, 512)
        if pid < 0:
            raise ValueError(err.value.decode())
        self._nobs.append(obs_op.n_obs())
        self._nparams.append(param_set.n_active())
        return pid

    def submit(self, problem: int, theta) -> int:
        """Queues one coefficient row; returns its ticket (raises
        BlockingIOError when the request ring is full)."""
        th = np.ascontiguousarray(theta, dtype=np.float64).ravel()
        if problem < 0 or problem >= len(self._nparams) or th.size != self._nparams[problem]:
            raise ValueError("theta must have the problem's n_parameters entries")
        err = C.create_string_buffer(512)
        t = self.lib.rq_server_submit(self._h, int(problem), th.ctypes.data_as(engine._dp),
                                      th.size, err, 512)
        if t == -2:
            raise BlockingIOError(err.value.decode())
        if t < 0:
            raise ValueError(err.value.decode())
        self._ticket_problem = getattr(self, "_ticket_problem", {})
        self._ticket_problem[t] = problem
        return int(t)

    def poll(self, ticket: int):
        """The ServedRow of a finished ticket, or None while it runs."""
        problem = self._ticket_problem[ticket]
        pred = np.full(self._nobs[problem], np.nan)
        row = RqServeRow()
        row.predictions = pred.ctypes.data_as(engine._dp)
        done = C.c_int32(0)
        err = C.create_string_buffer(512)
        rc = self.lib.rq_server_poll(self._h, int(ticket), C.byref(row), C.byref(done), err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        if not done.value:
            return None
        return ServedRow(engine.EvaluationStatus(row.status), row.loglik, pred, row.iterations,
                         row.final_worst_measure, row.contraction_rate_u, bool(row.stalled),
                         row.error.decode(), tuple(row.krylov_iterations))

    def wait(self, ticket: int, timeout: float = 600.0, interval: float = 1e-4):
        deadline = time.monotonic() + timeout
        while True:
            r = self.poll(ticket)
            if r is not None:
                return r
            if time.monotonic() > deadline:
                raise TimeoutError(f"ticket {ticket} not finished after {timeout} s")
            time.sleep(interval)

    def stats(self) -> dict:
        slots, nbytes = C.c_int32(), C.c_int64()
        sub, done = C.c_int64(), C.c_int64()
        self.lib.rq_server_stats(self._h, C.byref(slots), C.byref(nbytes), C.byref(sub),
                                 C.byref(done))
        return {"slots": slots.value, "slot_arena_bytes": nbytes.value,
                "submitted": sub.value, "completed": done.value}
