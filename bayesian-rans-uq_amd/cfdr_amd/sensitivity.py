"""Total sensitivities by warm-started central differences on the GPU
(include/rans_amd.h, rq_sensitivity_*).

``ParameterSensitivity`` mirrors the reference's class and binding for its
warm-FD path (rans-uq/include/ParameterSensitivity.hpp:68-133,
rans-uq/src/ParameterSensitivity.cpp:35-59,729-840,
rans-uq/python/bindings.cpp:1085-1089): ``solve_state(theta11)`` solves once
and keeps the converged state in device memory, and
``eta_jacobian_warm_fd(theta11, ...)`` re-solves the 22 perturbed coefficient
vectors from that state, all in one launch, and differences their
observations and log-likelihoods. Coefficient vectors are the eleven SST
coefficients in the reference's order.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import engine
from .config import FlowBoundaryConditions, ObservationOperator, SolverSettings, Vec3
from .mesh import Mesh


class RqWarmFdOut(C.Structure):
    _fields_ = [("d_obs_d_theta", engine._dp), ("log_lik_gradient", engine._dp),
                ("iterations", engine._ip), ("worst_measure", engine._dp),
                ("converged", engine._ip), ("steps", engine._dp), ("tolerances", engine._dp),
                ("positive_status", engine._ip), ("negative_status", engine._ip),
                ("n_residual_evals", engine._i64p)]


def _bind(lib):
    if getattr(lib, "_sens_bound", False):
        return lib
    lib.rq_sensitivity_create.argtypes = [C.POINTER(engine.RqMesh), C.POINTER(engine.RqObs),
                                          C.POINTER(engine.RqBc), C.c_double,
                                          C.POINTER(engine.RqSettings), engine._dp, C.c_char_p,
                                          C.c_size_t]
    lib.rq_sensitivity_create.restype = C.c_void_p
    lib.rq_sensitivity_solve_state.argtypes = [C.c_void_p, engine._dp, engine._ip, C.c_char_p,
                                               C.c_size_t]
    lib.rq_sensitivity_solve_state.restype = C.c_int
    lib.rq_sensitivity_has_state.argtypes = [C.c_void_p]
    lib.rq_sensitivity_has_state.restype = C.c_int32
    lib.rq_sensitivity_eta_jacobian_warm_fd.argtypes = [
        C.c_void_p, engine._dp, C.c_double, C.c_double, C.c_int32, C.c_double, engine._dp,
        C.POINTER(RqWarmFdOut), C.c_char_p, C.c_size_t]
    lib.rq_sensitivity_eta_jacobian_warm_fd.restype = C.c_int
    lib.rq_sensitivity_engine.argtypes = [C.c_void_p]
    lib.rq_sensitivity_engine.restype = C.c_void_p
    lib.rq_sensitivity_destroy.argtypes = [C.c_void_p]
    lib.rq_sensitivity_destroy.restype = None
    lib._sens_bound = True
    return lib


@dataclass
class TangentGradientResult:
    """The warm-FD fields of TangentGradientResult (ParameterSensitivity.hpp:24-36)."""

    d_obs_d_theta: np.ndarray        # n_obs x 11
    log_lik_gradient: np.ndarray     # 11
    krylov_iters: np.ndarray         # 11: SIMPLE iterations of the + and - solves
    krylov_rel_res: np.ndarray       # 11: the larger final convergence measure
    krylov_converged: np.ndarray     # 11
    finite_difference_steps: np.ndarray
    perturbed_solve_tolerances: np.ndarray
    positive_status: list
    negative_status: list
    n_residual_evals: int


class ParameterSensitivity:
    """Warm-FD total sensitivities with the base state on the GPU."""

    def __init__(self, mesh: Mesh, obs_op: ObservationOperator, bcs: FlowBoundaryConditions,
                 nu: float, settings: SolverSettings = None, u_init: Vec3 = Vec3(1.0, 0.0, 0.0),
                 p_init: float = 0.0, k_init: float = 1e-4, omega_init: float = 1.0):
        from .config import InferenceParameterSet
        self.lib = _bind(engine.load_library())
        # the descriptors of a ForwardBatch problem; the parameter set only
        # satisfies the constructor (the sensitivity uses all eleven)
        pr = engine.Problem(mesh, InferenceParameterSet.a1_betaStar(), obs_op, bcs, nu, settings,
                            u_init, p_init, k_init, omega_init, owner="ParameterSensitivity")
        self.n_obs = obs_op.n_obs()
        a = pr.args()
        err = C.create_string_buffer(512)
        h = self.lib.rq_sensitivity_create(a[0], a[2], a[3], a[4], a[5], a[6], err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h

    def close(self) -> None:
        """Releases the native object now (rq_sensitivity_destroy) instead of at garbage
        collection."""
        h = getattr(self, "_h", None)
        if h and getattr(self, "lib", None) is not None:
            self._h = None
            self.lib.rq_sensitivity_destroy(h)

    def __del__(self):
        self.close()

    @staticmethod
    def _theta(theta11) -> np.ndarray:
        t = np.ascontiguousarray(theta11, dtype=np.float64).ravel()
        if t.size != 11:
            raise ValueError("the coefficient vector needs eleven entries")
        return t

    def solve_state(self, theta11) -> engine.EvaluationStatus:
        t = self._theta(theta11)
        st = C.c_int32(0)
        err = C.create_string_buffer(512)
        rc = self.lib.rq_sensitivity_solve_state(self._h, t.ctypes.data_as(engine._dp),
                                                 C.byref(st), err, 512)
        engine.BatchBuffers.check(rc, err)
        return engine.EvaluationStatus(st.value)

    def has_state(self) -> bool:
        return bool(self.lib.rq_sensitivity_has_state(self._h))

    def eta_jacobian_warm_fd(self, theta, h_rel: float = 5e-4, h_floor: float = 1e-7,
                             warm_max_iter: int = 0, warm_tol: float = 0.0,
                             absolute_steps=()) -> TangentGradientResult:
        t = self._theta(theta)
        steps = None
        if len(absolute_steps):
            steps = np.ascontiguousarray(absolute_steps, dtype=np.float64).ravel()
            if steps.size != 11:
                raise ValueError("finite-difference steps require eleven positive finite entries")
        n = self.n_obs
        d = np.zeros((n, 11))
        g, worst, hs, tols = np.zeros(11), np.zeros(11), np.zeros(11), np.zeros(11)
        its, conv = np.zeros(11, np.int32), np.zeros(11, np.int32)
        ps, ms = np.zeros(11, np.int32), np.zeros(11, np.int32)
        evals = np.zeros(1, np.int64)
        out = RqWarmFdOut(d.ctypes.data_as(engine._dp), g.ctypes.data_as(engine._dp),
                          its.ctypes.data_as(engine._ip), worst.ctypes.data_as(engine._dp),
                          conv.ctypes.data_as(engine._ip), hs.ctypes.data_as(engine._dp),
                          tols.ctypes.data_as(engine._dp), ps.ctypes.data_as(engine._ip),
                          ms.ctypes.data_as(engine._ip), evals.ctypes.data_as(engine._i64p))
        err = C.create_string_buffer(512)
        rc = self.lib.rq_sensitivity_eta_jacobian_warm_fd(
            self._h, t.ctypes.data_as(engine._dp), float(h_rel), float(h_floor),
            int(warm_max_iter), float(warm_tol),
            None if steps is None else steps.ctypes.data_as(engine._dp), C.byref(out), err, 512)
        engine.BatchBuffers.check(rc, err)
        return TangentGradientResult(d, g, its, worst, conv, hs, tols,
                                     [engine.EvaluationStatus(int(v)) for v in ps],
                                     [engine.EvaluationStatus(int(v)) for v in ms], int(evals[0]))

    def last_kernel_ms(self) -> float:
        ms = C.c_double()
        e = self.lib.rq_sensitivity_engine(self._h)
        self.lib.rq_engine_last_timing(e, C.byref(ms), None, None, None, None)
        return ms.value
