"""ctypes binding of the gfx950 engine (include/rans_amd.h) and the
ForwardBatch front end.

``ForwardBatch`` mirrors the reference's batch forward model and its Python
binding (rans-uq/include/ForwardBatch.hpp:23-125,
rans-uq/python/bindings.cpp:830-916): the constructor deep-copies the
problem, ``evaluate(theta, workers)`` returns one row per coefficient vector
with the reference's status, log-likelihood and NaN conventions. All rows run
on the GPU in one persistent kernel launch. The engine library must be built
(``__graft_entry__.build()``); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
import os
from dataclasses import dataclass

import numpy as np

from .config import (BC_NEUMANN, PATCH_TYPES_ORDER, SCHEMES, SOLVERS, STATUS_NAMES,
                     FlowBoundaryConditions, InferenceParameterSet, ObservationOperator,
                     SolverSettings, Vec3, validate_coefficients)
from .mesh import Mesh

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librans_amd.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)


class RqMesh(C.Structure):
    _fields_ = [("n_cells", C.c_int32), ("n_internal", C.c_int32), ("n_faces", C.c_int32),
                ("n_patches", C.c_int32), ("owner", _ip), ("neighbor", _ip),
                ("face_center", _dp), ("face_normal", _dp), ("face_area", _dp),
                ("face_delta", _dp), ("face_d", _dp), ("face_weight", _dp), ("face_shift", _dp),
                ("cell_center", _dp), ("cell_volume", _dp), ("wall_distance", _dp),
                ("cell_face_ptr", _ip), ("cell_face_idx", _ip), ("patch_type", _ip),
                ("patch_face_ptr", _ip), ("patch_face_idx", _ip), ("cell_folded", _dp)]


class RqBc(C.Structure):
    _fields_ = [("u_type", _ip), ("p_type", _ip), ("k_type", _ip), ("w_type", _ip),
                ("u_val", _dp), ("p_val", _dp), ("k_val", _dp), ("w_val", _dp)]


class RqSettings(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("convergence_tol", C.c_double),
                ("divergence_limit", C.c_double), ("convergence_window", C.c_int32),
                ("monitor_interval", C.c_int32), ("alpha_u", C.c_double),
                ("alpha_p", C.c_double), ("alpha_k", C.c_double), ("alpha_omega", C.c_double),
                ("inner_iterations", C.c_int32), ("inner_tolerance", C.c_double),
                ("turb_start_iter", C.c_int32), ("turb_update_interval", C.c_int32),
                ("nut_floor_iters", C.c_int32), ("momentum_convection", C.c_int32),
                ("turbulence_convection", C.c_int32), ("body_force_x", C.c_double),
                ("body_force_y", C.c_double), ("k_min", C.c_double), ("k_max", C.c_double),
                ("omega_min", C.c_double), ("pressure_solver", C.c_int32),
                ("momentum_solver", C.c_int32), ("turbulence_solver", C.c_int32),
                ("sst_variant", C.c_int32), ("mixed_precision", C.c_int32),
                ("block_size", C.c_int32), ("profile", C.c_int32), ("grid_blocks", C.c_int32),
                ("amg_block_rows", C.c_int32), ("grid_plain_launch", C.c_int32)]


class RqParams(C.Structure):
    _fields_ = [("n_active", C.c_int32), ("active_indices", _ip), ("base", _dp),
                ("lower", _dp), ("upper", _dp)]


class RqObs(C.Structure):
    _fields_ = [("n_obs", C.c_int32), ("type", _ip), ("component", _ip), ("observed", _dp),
                ("variance", _dp), ("ref_velocity", _dp), ("ref_pressure", _dp),
                ("ref_density", _dp), ("cell", _ip), ("offx", _dp), ("offy", _dp),
                ("axx", _dp), ("axy", _dp), ("ayy", _dp), ("det", _dp), ("nb_ptr", _ip),
                ("nb_cell", _ip), ("nb_face", _ip), ("nb_dx", _dp), ("nb_dy", _dp),
                ("nb_w", _dp), ("face_a", _ip), ("face_b", _ip), ("t", _dp)]


class RqBatchOut(C.Structure):
    _fields_ = [("status", _ip), ("loglik", _dp), ("predictions", _dp), ("iterations", _ip),
                ("final_worst_measure", _dp), ("contraction_rate_u", _dp), ("stalled", _u8p),
                ("converged", _u8p), ("row_errors", C.c_char_p), ("row_error_stride", C.c_size_t),
                ("fields", _dp), ("krylov_iterations", _i64p), ("phase_cycles", _i64p)]


class RqTuneOut(C.Structure):
    _fields_ = [("n_candidates", C.c_int32), ("block_size", C.c_int32 * 16),
                ("grid_blocks", C.c_int32 * 16), ("amg_block_rows", C.c_int32 * 16),
                ("kernel_ms", C.c_double * 16), ("chosen_block_size", C.c_int32),
                ("chosen_grid_blocks", C.c_int32), ("chosen_amg_block_rows", C.c_int32)]


EXPORTS = ["rq_engine_amg_levels", "rq_engine_autotune", "rq_engine_set_graph_replay", "rq_engine_graph_launches", "rq_engine_create", "rq_engine_evaluate", "rq_engine_evaluate_noalloc", "rq_engine_n_observations",
           "rq_engine_n_parameters", "rq_engine_n_cells", "rq_engine_n_faces",
           "rq_engine_last_timing", "rq_engine_destroy", "rq_build_info",
           "rq_server_create", "rq_server_add_problem", "rq_server_submit", "rq_server_poll",
           "rq_server_stats", "rq_server_destroy",
           "rq_model_create", "rq_model_evaluate", "rq_model_evaluate_batch", "rq_model_engine",
           "rq_model_cache_size", "rq_model_cache_retained_bytes", "rq_model_cache_byte_budget",
           "rq_model_cache_set_byte_budget", "rq_model_cache_clear", "rq_model_destroy",
           "rq_warm_cache_create", "rq_warm_cache_store", "rq_warm_cache_nearest",
           "rq_warm_cache_size", "rq_warm_cache_retained_bytes", "rq_warm_cache_set_byte_budget",
           "rq_warm_cache_clear", "rq_warm_cache_entry_bytes", "rq_warm_cache_destroy",
           "rq_sensitivity_create", "rq_sensitivity_solve_state", "rq_sensitivity_has_state",
           "rq_sensitivity_eta_jacobian_warm_fd", "rq_sensitivity_engine",
           "rq_sensitivity_destroy"]

_lib = None


class EngineMissing(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Load the engine library; raises EngineMissing when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineMissing(f"engine library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    lib.rq_engine_create.argtypes = [C.POINTER(RqMesh), C.POINTER(RqParams), C.POINTER(RqObs),
                                     C.POINTER(RqBc), C.c_double, C.POINTER(RqSettings), _dp,
                                     C.c_char_p, C.c_size_t]
    lib.rq_engine_create.restype = C.c_void_p
    lib.rq_engine_evaluate.argtypes = [C.c_void_p, _dp, C.c_int64, C.c_int32, C.c_int32,
                                       C.POINTER(RqBatchOut), C.c_char_p, C.c_size_t]
    lib.rq_engine_evaluate.restype = C.c_int
    lib.rq_engine_evaluate_noalloc.argtypes = lib.rq_engine_evaluate.argtypes
    lib.rq_engine_evaluate_noalloc.restype = C.c_int
    for name in ("rq_engine_n_observations", "rq_engine_n_parameters", "rq_engine_n_cells",
                 "rq_engine_n_faces"):
        getattr(lib, name).argtypes = [C.c_void_p]
        getattr(lib, name).restype = C.c_int32
    lib.rq_engine_last_timing.argtypes = [C.c_void_p, _dp, _ip, _ip, _i64p, _ip]
    lib.rq_engine_last_timing.restype = C.c_int
    lib.rq_engine_destroy.argtypes = [C.c_void_p]
    lib.rq_engine_destroy.restype = None
    lib.rq_engine_autotune.argtypes = [C.c_void_p, _dp, C.c_int64, C.c_int32, C.c_int32,
                                       C.POINTER(RqTuneOut), C.c_char_p, C.c_size_t]
    lib.rq_engine_autotune.restype = C.c_int
    lib.rq_engine_set_graph_replay.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_size_t]
    lib.rq_engine_set_graph_replay.restype = C.c_int
    lib.rq_engine_amg_levels.argtypes = [C.c_void_p, _ip, _i64p, C.c_int32]
    lib.rq_engine_amg_levels.restype = C.c_int32
    lib.rq_engine_graph_launches.argtypes = [C.c_void_p]
    lib.rq_engine_graph_launches.restype = C.c_int64
    lib.rq_build_info.argtypes = []
    lib.rq_build_info.restype = C.c_char_p
    _lib = lib
    return lib


class _Keep:
    def __init__(self):
        self.items = []

    def d(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.items.append(a)
        return a.ctypes.data_as(_dp)

    def i(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        self.items.append(a)
        return a.ctypes.data_as(_ip)


def settings_struct(s: SolverSettings, mixed_precision: bool = False, block_size: int = 0,
                    profile: bool = False, grid_blocks: int = 0, amg_block_rows: int = 0,
                    grid_plain_launch: bool = False):
    r = RqSettings()
    r.max_iterations, r.convergence_tol, r.divergence_limit = s.max_iterations, \
        s.convergence_tol, s.divergence_limit
    r.convergence_window, r.monitor_interval = s.convergence_window, s.monitor_interval
    r.alpha_u, r.alpha_p, r.alpha_k, r.alpha_omega = s.alpha_u, s.alpha_p, s.alpha_k, s.alpha_omega
    r.inner_iterations, r.inner_tolerance = s.inner_iterations, s.inner_tolerance
    r.turb_start_iter, r.turb_update_interval = s.turb_start_iter, s.turb_update_interval
    r.nut_floor_iters = s.nut_floor_iters
    r.momentum_convection = SCHEMES[s.momentum_convection]
    r.turbulence_convection = SCHEMES[s.turbulence_convection]
    r.body_force_x, r.body_force_y = s.body_force.x, s.body_force.y
    r.k_min, r.k_max, r.omega_min = s.k_min, s.k_max, s.omega_min
    r.pressure_solver = SOLVERS[s.pressure_solver]
    r.momentum_solver = SOLVERS[s.momentum_solver]
    r.turbulence_solver = SOLVERS[s.turbulence_solver]
    r.sst_variant = s.sst_variant
    r.mixed_precision = 1 if mixed_precision else 0
    r.block_size = block_size
    r.profile = 1 if profile else 0
    r.grid_blocks = grid_blocks
    r.amg_block_rows = amg_block_rows
    r.grid_plain_launch = 1 if grid_plain_launch else 0
    return r


def planar_problem(mesh: Mesh, bcs: FlowBoundaryConditions):
    """The mesh and boundary conditions the engine solves on: a planar mesh
    as it is; a one-layer extruded mesh (an OpenFOAM 2D case) with its
    front and back faces folded out (Mesh.fold_empty_faces) and its "empty"
    patches' conditions dropped. Those must stay zero-gradient in every
    field, as channelDefaults leaves a patch of another type
    (BoundaryCondition.hpp:160-166): only then is a folded face the
    zero-gradient term the engine adds. Returns (mesh, bcs, folded or None)."""
    planar, kept, folded = mesh.fold_empty_faces()
    if folded is None:
        return mesh, bcs, None
    for i in range(mesh.n_patches()):
        if i in kept:
            continue
        for lst in (bcs.velocity, bcs.pressure, bcs.k, bcs.omega):
            if lst[i].type != BC_NEUMANN:
                raise ValueError(f"patch {mesh.patches[i].name!r} of type 'empty' needs "
                                 "zero-gradient conditions in every field")
    pick = lambda lst: [lst[i] for i in kept]  # noqa: E731
    return planar, FlowBoundaryConditions(pick(bcs.velocity), pick(bcs.pressure),
                                          pick(bcs.k), pick(bcs.omega)), folded


def mesh_struct(m: Mesh, keep: _Keep, folded: np.ndarray = None):
    r = RqMesh()
    r.n_cells, r.n_internal, r.n_faces, r.n_patches = m.n_cells, m.n_internal, m.n_faces, \
        m.n_patches()
    r.owner, r.neighbor = keep.i(m.owner), keep.i(m.neighbor)
    r.face_center, r.face_normal = keep.d(m.face_center), keep.d(m.face_normal)
    r.face_area, r.face_delta = keep.d(m.face_area), keep.d(m.face_delta)
    r.face_d, r.face_weight, r.face_shift = keep.d(m.face_d), keep.d(m.face_weight), \
        keep.d(m.face_shift)
    r.cell_center, r.cell_volume = keep.d(m.cell_center), keep.d(m.cell_volume)
    r.wall_distance = keep.d(m.wall_distance())
    r.cell_face_ptr, r.cell_face_idx = keep.i(m.cell_face_ptr), keep.i(m.cell_face_idx)
    r.patch_type = keep.i([PATCH_TYPES_ORDER.get(p.type, 4) for p in m.patches] or [0])
    ptr = np.zeros(m.n_patches() + 1, np.int32)
    ptr[1:] = np.cumsum([len(p.faces) for p in m.patches])
    r.patch_face_ptr = keep.i(ptr)
    r.patch_face_idx = keep.i(np.concatenate([p.faces for p in m.patches])
                              if m.patches else np.zeros(1, np.int32))
    r.cell_folded = keep.d(folded) if folded is not None else None
    return r


def obs_struct(P, keep: _Keep):
    r = RqObs()
    r.n_obs = P.n_obs
    r.type, r.component = keep.i(P.type), keep.i(P.component)
    r.observed, r.variance = keep.d(P.observed), keep.d(P.variance)
    r.ref_velocity, r.ref_pressure, r.ref_density = keep.d(P.ref_velocity), \
        keep.d(P.ref_pressure), keep.d(P.ref_density)
    r.cell, r.offx, r.offy = keep.i(P.cell), keep.d(P.offx), keep.d(P.offy)
    r.axx, r.axy, r.ayy, r.det = keep.d(P.axx), keep.d(P.axy), keep.d(P.ayy), keep.d(P.det)
    r.nb_ptr, r.nb_cell, r.nb_face = keep.i(P.nb_ptr), keep.i(P.nb_cell), keep.i(P.nb_face)
    r.nb_dx, r.nb_dy, r.nb_w = keep.d(P.nb_dx), keep.d(P.nb_dy), keep.d(P.nb_w)
    r.face_a, r.face_b, r.t = keep.i(P.face_a), keep.i(P.face_b), keep.d(P.t)
    return r


class EvaluationStatus(enum.IntEnum):
    """EvaluationStatus (core/include/EvaluationTypes.hpp:6) with the names
    of the reference's Python binding (rans-uq/python/bindings.cpp:464-469)."""

    Converged = 0
    Unconverged = 1
    DivergenceDetected = 2
    InvalidParameters = 3
    Unknown = 4


assert [s.name for s in EvaluationStatus] == STATUS_NAMES


@dataclass
class EvaluationResult:
    """EvaluationResult (core/include/EvaluationTypes.hpp:24-29)."""

    status: EvaluationStatus
    log_lik: float
    predictions: np.ndarray
    simple_iters: int

    @property
    def status_name(self) -> str:
        return self.status.name


@dataclass
class ForwardBatchRow:
    """ForwardBatchRow (rans-uq/include/ForwardBatch.hpp:8-14)."""

    result: EvaluationResult
    final_worst_measure: float
    contraction_rate_u: float
    stalled: bool
    error: str


@dataclass
class ForwardBatchResult:
    """ForwardBatchResult with the array views of the reference's binding
    (rans-uq/python/bindings.cpp:830-887); every array is owned. `statuses`
    is a list of EvaluationStatus as in the binding; `status_codes` holds the
    same values as an int array."""

    n_observations: int
    rows: list
    predictions: np.ndarray
    log_likelihoods: np.ndarray
    statuses: list
    iteration_counts: np.ndarray
    errors: list
    convergence_mask: np.ndarray
    final_worst_measures: np.ndarray
    contraction_rates: np.ndarray
    stalled: np.ndarray
    krylov_iterations: np.ndarray
    fields: np.ndarray = None
    kernel_ms: float = 0.0
    phase_cycles: np.ndarray = None

    @property
    def converged(self) -> np.ndarray:
        return self.convergence_mask

    @property
    def status_codes(self) -> np.ndarray:
        return np.asarray([int(s) for s in self.statuses], dtype=np.int32)


class Problem:
    """The problem inputs flattened into the C-ABI descriptors (validated as
    the reference's constructors do, ForwardBatch.hpp:34-50)."""

    def __init__(self, mesh: Mesh, param_set: InferenceParameterSet,
                 obs_op: ObservationOperator, bcs: FlowBoundaryConditions, nu: float,
                 settings: SolverSettings = None, u_init: Vec3 = Vec3(1.0, 0.0, 0.0),
                 p_init: float = 0.0, k_init: float = 1e-4, omega_init: float = 1.0,
                 mixed_precision: bool = False, block_size: int = 0, profile: bool = False,
                 grid_blocks: int = 0, owner: str = "ForwardBatch", amg_block_rows: int = 0,
                 grid_plain_launch: bool = False):
        settings = settings if settings is not None else SolverSettings()
        settings.validate()
        param_set.validate()
        bcs.validate(mesh)
        if not mesh.wall_distance_valid():
            raise ValueError(f"{owner} needs compute_wall_distance before construction")
        mesh, bcs, folded = planar_problem(mesh, bcs)
        if not math.isfinite(nu) or nu <= 0:
            raise ValueError("SIMPLE kinematic viscosity must be finite and positive")
        if isinstance(u_init, (tuple, list)):
            u_init = Vec3(*u_init)
        if u_init.z != 0.0:
            raise ValueError("planar solver: initial velocity must have zero z component")
        prepared = obs_op.prepare(mesh)
        self.keep = keep = _Keep()
        self.m = mesh_struct(mesh, keep, folded)
        b = self.b = RqBc()
        ut, uv, pt, pv, kt, kv, wt, wv = bcs.arrays()
        b.u_type, b.p_type, b.k_type, b.w_type = keep.i(ut), keep.i(pt), keep.i(kt), keep.i(wt)
        b.u_val, b.p_val, b.k_val, b.w_val = keep.d(uv), keep.d(pv), keep.d(kv), keep.d(wv)
        p = self.p = RqParams()
        p.n_active = param_set.n_active()
        p.active_indices = keep.i(param_set.active_indices)
        p.base = keep.d(param_set.base)
        p.lower, p.upper = keep.d(param_set.lower_bounds()), keep.d(param_set.upper_bounds())
        self.o = obs_struct(prepared, keep)
        self.s = settings_struct(settings, mixed_precision, block_size, profile, grid_blocks,
                                 amg_block_rows, grid_plain_launch)
        self.init = np.asarray([u_init.x, u_init.y, p_init, k_init, omega_init], float)
        self.nu = float(nu)

    def args(self):
        return (C.byref(self.m), C.byref(self.p), C.byref(self.o), C.byref(self.b), self.nu,
                C.byref(self.s), self.init.ctypes.data_as(_dp))


class BatchBuffers:
    """Host arrays behind one rq_batch_out of n rows, and their decoding into
    a ForwardBatchResult."""

    stride = 256

    def __init__(self, n: int, nobs: int, n_cells: int, n_faces: int, fields: bool = False):
        self.n, self.nobs = n, nobs
        self.status = np.zeros(n, np.int32)
        self.loglik = np.zeros(n)
        self.pred = np.full((n, max(nobs, 0)), np.nan)
        self.iters = np.zeros(n, np.int32)
        self.worst = np.zeros(n)
        self.rate = np.zeros(n)
        self.stalled = np.zeros(n, np.uint8)
        self.conv = np.zeros(n, np.uint8)
        self.kry = np.zeros((n, 5), np.int64)
        self.cyc = np.zeros((n, 16), np.int64)
        self.errs = C.create_string_buffer(max(1, n * self.stride))
        self.fld = np.zeros((n, 10 * n_cells + n_faces)) if fields else None
        out = self.out = RqBatchOut()
        out.status = self.status.ctypes.data_as(_ip)
        out.loglik = self.loglik.ctypes.data_as(_dp)
        out.predictions = self.pred.ctypes.data_as(_dp)
        out.iterations = self.iters.ctypes.data_as(_ip)
        out.final_worst_measure = self.worst.ctypes.data_as(_dp)
        out.contraction_rate_u = self.rate.ctypes.data_as(_dp)
        out.stalled = self.stalled.ctypes.data_as(_u8p)
        out.converged = self.conv.ctypes.data_as(_u8p)
        out.row_errors = C.cast(self.errs, C.c_char_p)
        out.row_error_stride = self.stride
        out.fields = self.fld.ctypes.data_as(_dp) if fields else None
        out.krylov_iterations = self.kry.ctypes.data_as(_i64p)
        out.phase_cycles = self.cyc.ctypes.data_as(_i64p)

    @staticmethod
    def check(rc: int, err) -> None:
        if rc == 1:
            raise ValueError(err.value.decode())
        if rc == 3:
            raise MemoryError(err.value.decode())
        if rc != 0:
            raise RuntimeError(err.value.decode())

    def result(self, kernel_ms: float = 0.0) -> "ForwardBatchResult":
        n, st = self.n, self.stride
        raw = self.errs.raw
        errors = [raw[i * st:(i + 1) * st].split(b"\0", 1)[0].decode() for i in range(n)]
        statuses = [EvaluationStatus(int(v)) for v in self.status]
        rows = []
        for i in range(n):
            res = EvaluationResult(statuses[i], float(self.loglik[i]), self.pred[i].copy(),
                                   int(self.iters[i]))
            rows.append(ForwardBatchRow(res, float(self.worst[i]), float(self.rate[i]),
                                        bool(self.stalled[i]), errors[i]))
        return ForwardBatchResult(self.nobs, rows, self.pred, self.loglik, statuses, self.iters,
                                  errors, self.conv.astype(bool), self.worst, self.rate,
                                  self.stalled.astype(bool), self.kry, self.fld, kernel_ms,
                                  self.cyc)


class ForwardBatch:
    """Synchronous coefficient batch on one GPU (ForwardBatch.hpp:23-125)."""

    def __init__(self, mesh: Mesh, param_set: InferenceParameterSet,
                 obs_op: ObservationOperator, bcs: FlowBoundaryConditions, nu: float,
                 settings: SolverSettings = None, u_init: Vec3 = Vec3(1.0, 0.0, 0.0),
                 p_init: float = 0.0, k_init: float = 1e-4, omega_init: float = 1.0,
                 mixed_precision: bool = False, block_size: int = 0, profile: bool = False,
                 grid_blocks: int = 0, amg_block_rows: int = 0, grid_plain_launch: bool = False):
        self.lib = load_library()
        pr = Problem(mesh, param_set, obs_op, bcs, nu, settings, u_init, p_init, k_init,
                     omega_init, mixed_precision, block_size, profile, grid_blocks, "ForwardBatch",
                     amg_block_rows, grid_plain_launch)
        self.param_set = param_set
        self.n_obs = obs_op.n_obs()
        self.n_cells = mesh.n_cells
        self.n_faces = int(pr.m.n_faces)  # in-plane faces (front/back faces are folded)
        err = C.create_string_buffer(512)
        h = self.lib.rq_engine_create(*pr.args(), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h

    def close(self) -> None:
        """Releases the engine's device memory, stream and captured graph now
        (rq_engine_destroy) instead of at garbage collection."""
        h = getattr(self, "_h", None)
        if h and getattr(self, "lib", None) is not None:
            self._h = None
            self.lib.rq_engine_destroy(h)

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def n_parameters(self) -> int:
        return self.param_set.n_active()

    def n_observations(self) -> int:
        return self.n_obs

    def last_timing(self):
        ms, bs, lds, nbytes = C.c_double(), C.c_int32(), C.c_int32(), C.c_int64()
        gb = C.c_int32()
        self.lib.rq_engine_last_timing(self._h, C.byref(ms), C.byref(bs), C.byref(lds),
                                       C.byref(nbytes), C.byref(gb))
        return {"kernel_ms": ms.value, "block_size": bs.value, "arena_in_lds": bool(lds.value),
                "arena_bytes": nbytes.value, "grid_blocks": gb.value}

    def autotune(self, theta, iterations: int = 200) -> dict:
        """Times every candidate launch shape on these rows (each solve capped
        at `iterations` outer iterations) and keeps the fastest for later
        evaluations (rq_engine_autotune)."""
        theta = np.ascontiguousarray(np.atleast_2d(np.asarray(theta, dtype=np.float64)))
        if theta.shape[1] != self.n_parameters():
            raise ValueError("theta must have shape (n_rows, n_parameters)")
        out = RqTuneOut()
        err = C.create_string_buffer(512)
        rc = self.lib.rq_engine_autotune(self._h, theta.ctypes.data_as(_dp), theta.shape[0],
                                         self.n_parameters(), int(iterations), C.byref(out),
                                         err, 512)
        BatchBuffers.check(rc, err)
        cands = [{"block_size": out.block_size[i], "grid_blocks": out.grid_blocks[i],
                  "amg_block_rows": out.amg_block_rows[i], "kernel_ms": out.kernel_ms[i]}
                 for i in range(out.n_candidates)]
        return {"candidates": cands, "block_size": out.chosen_block_size,
                "grid_blocks": out.chosen_grid_blocks,
                "amg_block_rows": out.chosen_amg_block_rows}

    def set_graph_replay(self, enable: bool = True) -> None:
        """Capture the batch launch into a HIP graph once and replay it while
        the row count and launch shape stay the same (rq_engine_set_graph_replay)."""
        err = C.create_string_buffer(512)
        BatchBuffers.check(self.lib.rq_engine_set_graph_replay(self._h, 1 if enable else 0, err,
                                                               512), err)

    def amg_levels(self) -> list:
        """(rows, CSR entries) of each level of the pressure multigrid
        hierarchy (rq_engine_amg_levels); empty without AMG."""
        rows = np.zeros(16, np.int32)
        nnz = np.zeros(16, np.int64)
        n = self.lib.rq_engine_amg_levels(self._h, rows.ctypes.data_as(_ip),
                                          nnz.ctypes.data_as(_i64p), 16)
        return [(int(rows[l]), int(nnz[l])) for l in range(max(n, 0))]

    def graph_launches(self) -> int:
        return int(self.lib.rq_engine_graph_launches(self._h))

    def evaluate(self, theta, workers: int = 1, fields: bool = False) -> ForwardBatchResult:
        theta = np.asarray(theta, dtype=np.float64)
        if theta.ndim == 1 and theta.size == 0:
            theta = theta.reshape(0, self.n_parameters())
        if theta.ndim != 2 or theta.shape[1] != self.n_parameters():
            raise ValueError("theta must have shape (n_rows, n_parameters)")
        if workers < 1:
            raise ValueError("workers must be positive")
        theta = np.ascontiguousarray(theta)
        bb = BatchBuffers(theta.shape[0], self.n_obs, self.n_cells, self.n_faces, fields)
        err = C.create_string_buffer(512)
        rc = self.lib.rq_engine_evaluate(self._h, theta.ctypes.data_as(_dp), theta.shape[0],
                                         self.n_parameters(), int(workers), C.byref(bb.out), err,
                                         512)
        BatchBuffers.check(rc, err)
        return bb.result(self.last_timing()["kernel_ms"])

    def evaluate_noalloc(self, theta, workers: int = 1) -> ForwardBatchResult:
        """The same evaluation through rq_engine_evaluate_noalloc: every
        per-call device buffer comes from the library's fixed scratch region
        (no device allocation in the call); no fields are returned."""
        theta = np.asarray(theta, dtype=np.float64)
        if theta.ndim == 1 and theta.size == 0:
            theta = theta.reshape(0, self.n_parameters())
        if theta.ndim != 2 or theta.shape[1] != self.n_parameters():
            raise ValueError("theta must have shape (n_rows, n_parameters)")
        theta = np.ascontiguousarray(theta)
        bb = BatchBuffers(theta.shape[0], self.n_obs, self.n_cells, self.n_faces, False)
        err = C.create_string_buffer(512)
        rc = self.lib.rq_engine_evaluate_noalloc(self._h, theta.ctypes.data_as(_dp),
                                                 theta.shape[0], self.n_parameters(), int(workers),
                                                 C.byref(bb.out), err, 512)
        BatchBuffers.check(rc, err)
        return bb.result(self.last_timing()["kernel_ms"])


def build_info() -> str:
    return load_library().rq_build_info().decode()
