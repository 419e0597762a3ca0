"""ctypes bridge to the CPU FP64 restatement (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrMesh(C.Structure):
    _fields_ = [("n_cells", C.c_int), ("n_internal", C.c_int), ("n_faces", C.c_int),
                ("n_patches", C.c_int), ("owner", _ip), ("neighbor", _ip),
                ("fcx", _dp), ("fcy", _dp), ("fnx", _dp), ("fny", _dp), ("farea", _dp),
                ("fdelta", _dp), ("fdx", _dp), ("fdy", _dp), ("fweight", _dp),
                ("fshx", _dp), ("fshy", _dp), ("ccx", _dp), ("ccy", _dp), ("vol", _dp),
                ("wall_dist", _dp), ("cell_face_ptr", _ip), ("cell_face_idx", _ip),
                ("patch_type", _ip), ("patch_face_ptr", _ip), ("patch_face_idx", _ip),
                ("cell_folded", _dp)]


class OrBc(C.Structure):
    _fields_ = [("u_type", _ip), ("p_type", _ip), ("k_type", _ip), ("w_type", _ip),
                ("u_val", _dp), ("p_val", _dp), ("k_val", _dp), ("w_val", _dp)]


class OrSettings(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("convergence_tol", C.c_double),
                ("divergence_limit", C.c_double), ("convergence_window", C.c_int),
                ("monitor_interval", C.c_int), ("alpha_u", C.c_double), ("alpha_p", C.c_double),
                ("alpha_k", C.c_double), ("alpha_omega", C.c_double),
                ("inner_iterations", C.c_int), ("inner_tolerance", C.c_double),
                ("turb_start_iter", C.c_int), ("turb_update_interval", C.c_int),
                ("nut_floor_iters", C.c_int), ("momentum_convection", C.c_int),
                ("turbulence_convection", C.c_int), ("body_force_x", C.c_double),
                ("body_force_y", C.c_double), ("k_min", C.c_double), ("k_max", C.c_double),
                ("omega_min", C.c_double), ("pressure_solver", C.c_int),
                ("momentum_solver", C.c_int), ("turbulence_solver", C.c_int),
                ("sst_variant", C.c_int)]


class OrObs(C.Structure):
    _fields_ = [("n_obs", C.c_int), ("type", _ip), ("component", _ip), ("observed", _dp),
                ("variance", _dp), ("ref_velocity", _dp), ("ref_pressure", _dp),
                ("ref_density", _dp), ("cell", _ip), ("offx", _dp), ("offy", _dp),
                ("axx", _dp), ("axy", _dp), ("ayy", _dp), ("det", _dp), ("nb_ptr", _ip),
                ("nb_cell", _ip), ("nb_face", _ip), ("nb_dx", _dp), ("nb_dy", _dp),
                ("nb_w", _dp), ("face_a", _ip), ("face_b", _ip), ("t", _dp)]


class OrResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("stalled", C.c_int),
                ("converged", C.c_int), ("diverged", C.c_int), ("worst", C.c_double),
                ("rate", C.c_double), ("loglik", C.c_double), ("predictions", _dp),
                ("Ux", _dp), ("Uy", _dp), ("p", _dp), ("k", _dp), ("omega", _dp),
                ("nuT", _dp), ("F1", _dp), ("F2", _dp), ("Pk", _dp), ("CDkw", _dp),
                ("flux", _dp), ("krylov_iters", C.c_longlong * 5), ("error", C.c_char * 256)]


_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "build/liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(LIB_PATH)
        _lib.or_evaluate.argtypes = [C.POINTER(OrMesh), C.POINTER(OrBc), C.POINTER(OrSettings),
                                     C.c_double, _dp, _dp, C.POINTER(OrObs), C.POINTER(OrResult)]
        _lib.or_evaluate.restype = C.c_int
        _lib.or_evaluate_ex.argtypes = [C.POINTER(OrMesh), C.POINTER(OrBc),
                                        C.POINTER(OrSettings), C.c_double, _dp, _dp,
                                        C.POINTER(OrObs), _dp, _dp, C.c_int, C.c_double,
                                        C.c_int, C.POINTER(OrResult)]
        _lib.or_evaluate_ex.restype = C.c_int
        for name, nargs in (("or_sst_f1", 6), ("or_sst_f2", 5), ("or_sst_pk", 5),
                            ("or_sst_pomega", 5)):
            f = getattr(_lib, name)
            f.argtypes = [_dp] + [C.c_double] * (nargs - 1)
            f.restype = C.c_double
        _lib.or_sst_nut.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        _lib.or_sst_nut.restype = C.c_double
        _lib.or_linear_solve.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp,
                                         _dp, C.c_int, C.c_double, _ip, _dp, _dp]
        _lib.or_linear_solve.restype = C.c_int
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class _Keep:
    """Holds numpy buffers alive while C structs point at them."""

    def __init__(self):
        self.items = []

    def d(self, a):
        a = _d(a)
        self.items.append(a)
        return a.ctypes.data_as(_dp)

    def i(self, a):
        a = _i(a)
        self.items.append(a)
        return a.ctypes.data_as(_ip)


def planar(mesh, bcs):
    """A one-layer extruded mesh (an OpenFOAM 2D case) with its front and
    back faces folded out and its "empty" patches' conditions dropped, as
    the reference's solve sees them (zero-gradient faces without flux);
    (mesh, bcs, None) for a planar mesh."""
    from cfdr_amd.config import BC_NEUMANN, FlowBoundaryConditions
    flat, kept, folded = mesh.fold_empty_faces()
    if folded is None:
        return mesh, bcs, None
    for i in range(mesh.n_patches()):
        if i not in kept and any(lst[i].type != BC_NEUMANN
                                 for lst in (bcs.velocity, bcs.pressure, bcs.k, bcs.omega)):
            raise ValueError("an 'empty' patch needs zero-gradient conditions in every field")
    pick = lambda lst: [lst[i] for i in kept]  # noqa: E731
    return flat, FlowBoundaryConditions(pick(bcs.velocity), pick(bcs.pressure), pick(bcs.k),
                                        pick(bcs.omega)), folded


def marshal(mesh, bcs, settings, prepared, keep: _Keep, folded=None):
    from cfdr_amd.config import PATCH_TYPES_ORDER, SCHEMES, SOLVERS  # noqa: F401
    m = OrMesh()
    m.n_cells, m.n_internal, m.n_faces = mesh.n_cells, mesh.n_internal, mesh.n_faces
    m.n_patches = mesh.n_patches()
    m.owner, m.neighbor = keep.i(mesh.owner), keep.i(mesh.neighbor)
    fc, fn, fd, fs = mesh.face_center, mesh.face_normal, mesh.face_d, mesh.face_shift
    m.fcx, m.fcy = keep.d(fc[:, 0]), keep.d(fc[:, 1])
    m.fnx, m.fny = keep.d(fn[:, 0]), keep.d(fn[:, 1])
    m.farea, m.fdelta = keep.d(mesh.face_area), keep.d(mesh.face_delta)
    m.fdx, m.fdy = keep.d(fd[:, 0]), keep.d(fd[:, 1])
    m.fweight = keep.d(mesh.face_weight)
    m.fshx, m.fshy = keep.d(fs[:, 0]), keep.d(fs[:, 1])
    m.ccx, m.ccy = keep.d(mesh.cell_center[:, 0]), keep.d(mesh.cell_center[:, 1])
    m.vol, m.wall_dist = keep.d(mesh.cell_volume), keep.d(mesh.wall_distance())
    m.cell_face_ptr, m.cell_face_idx = keep.i(mesh.cell_face_ptr), keep.i(mesh.cell_face_idx)
    ptypes = [PATCH_TYPES_ORDER.get(p.type, 4) for p in mesh.patches]
    m.patch_type = keep.i(ptypes)
    pptr = np.zeros(len(mesh.patches) + 1, np.int32)
    pptr[1:] = np.cumsum([len(p.faces) for p in mesh.patches])
    m.patch_face_ptr = keep.i(pptr)
    m.patch_face_idx = keep.i(np.concatenate([p.faces for p in mesh.patches])
                              if mesh.patches else np.zeros(1, np.int32))
    m.cell_folded = keep.d(folded) if folded is not None else None
    b = OrBc()
    ut, uv, pt, pv, kt, kv, wt, wv = bcs.arrays()
    b.u_type, b.p_type, b.k_type, b.w_type = keep.i(ut), keep.i(pt), keep.i(kt), keep.i(wt)
    b.u_val, b.p_val, b.k_val, b.w_val = keep.d(uv), keep.d(pv), keep.d(kv), keep.d(wv)
    s = OrSettings()
    st = settings
    s.max_iterations, s.convergence_tol, s.divergence_limit = st.max_iterations, \
        st.convergence_tol, st.divergence_limit
    s.convergence_window, s.monitor_interval = st.convergence_window, st.monitor_interval
    s.alpha_u, s.alpha_p, s.alpha_k, s.alpha_omega = st.alpha_u, st.alpha_p, st.alpha_k, \
        st.alpha_omega
    s.inner_iterations, s.inner_tolerance = st.inner_iterations, st.inner_tolerance
    s.turb_start_iter, s.turb_update_interval = st.turb_start_iter, st.turb_update_interval
    s.nut_floor_iters = st.nut_floor_iters
    s.momentum_convection = SCHEMES[st.momentum_convection]
    s.turbulence_convection = SCHEMES[st.turbulence_convection]
    s.body_force_x, s.body_force_y = st.body_force.x, st.body_force.y
    s.k_min, s.k_max, s.omega_min = st.k_min, st.k_max, st.omega_min
    s.pressure_solver = SOLVERS[st.pressure_solver]
    s.momentum_solver = SOLVERS[st.momentum_solver]
    s.turbulence_solver = SOLVERS[st.turbulence_solver]
    s.sst_variant = st.sst_variant
    o = OrObs()
    P = prepared
    o.n_obs = P.n_obs
    o.type, o.component = keep.i(P.type), keep.i(P.component)
    o.observed, o.variance = keep.d(P.observed), keep.d(P.variance)
    o.ref_velocity, o.ref_pressure, o.ref_density = keep.d(P.ref_velocity), \
        keep.d(P.ref_pressure), keep.d(P.ref_density)
    o.cell, o.offx, o.offy = keep.i(P.cell), keep.d(P.offx), keep.d(P.offy)
    o.axx, o.axy, o.ayy, o.det = keep.d(P.axx), keep.d(P.axy), keep.d(P.ayy), keep.d(P.det)
    o.nb_ptr, o.nb_cell, o.nb_face = keep.i(P.nb_ptr), keep.i(P.nb_cell), keep.i(P.nb_face)
    o.nb_dx, o.nb_dy, o.nb_w = keep.d(P.nb_dx), keep.d(P.nb_dy), keep.d(P.nb_w)
    o.face_a, o.face_b, o.t = keep.i(P.face_a), keep.i(P.face_b), keep.d(P.t)
    return m, b, s, o


class OracleProblem:
    """One marshalled problem; ``evaluate`` runs one coefficient vector."""

    def __init__(self, mesh, bcs, settings, obs_op, nu, init):
        self.keep = _Keep()
        mesh, bcs, folded = planar(mesh, bcs)
        self.mesh = mesh
        self.prepared = obs_op.prepare(mesh)
        self.structs = marshal(mesh, bcs, settings, self.prepared, self.keep, folded)
        self.nu = float(nu)
        self.init = _d(init)

    def state_size(self) -> int:
        """Doubles of a flow state (warm start and save layout)."""
        nb = self.mesh.n_faces - self.mesh.n_internal
        return 10 * self.mesh.n_cells + 5 * max(nb, 1) + 2

    def evaluate(self, coeffs, fields: bool = False, warm=None, save: bool = False,
                 flags: int = 0, tol: float = 0.0, max_iter: int = 0):
        """One solve; warm: a state to start from; save: return the final
        state as "state"; flags: OR_FLAG_* bits; tol / max_iter > 0 override
        the settings."""
        L = lib()
        m, b, s, o = self.structs
        r = OrResult()
        pred = np.full(self.prepared.n_obs, np.nan)
        r.predictions = pred.ctypes.data_as(_dp)
        out = {}
        if fields:
            for name in ("Ux", "Uy", "p", "k", "omega", "nuT", "F1", "F2", "Pk", "CDkw"):
                out[name] = np.zeros(self.mesh.n_cells)
                setattr(r, name, out[name].ctypes.data_as(_dp))
            out["flux"] = np.zeros(self.mesh.n_faces)
            r.flux = out["flux"].ctypes.data_as(_dp)
        c = _d(coeffs)
        w = None if warm is None else _d(warm)
        st = np.zeros(self.state_size()) if save else None
        rc = L.or_evaluate_ex(C.byref(m), C.byref(b), C.byref(s), self.nu,
                              c.ctypes.data_as(_dp), self.init.ctypes.data_as(_dp), C.byref(o),
                              None if w is None else w.ctypes.data_as(_dp),
                              None if st is None else st.ctypes.data_as(_dp), int(flags),
                              float(tol), int(max_iter), C.byref(r))
        res = {"rc": rc, "status": r.status, "iters": r.iters, "stalled": bool(r.stalled),
               "converged": bool(r.converged), "diverged": bool(r.diverged),
               "worst": r.worst, "rate": r.rate, "loglik": r.loglik, "predictions": pred,
               "krylov_iters": list(r.krylov_iters), "error": r.error.decode()}
        res.update(out)
        if save:
            res["state"] = st
        return res


OR_FLAG_NO_MONITOR = 1
OR_FLAG_WARM_BCS = 2
ALL11_LOWER = (0.1, 0.1, 0.01, 0.1, 0.1, 0.1, 0.01, 0.1, 0.01, 0.1, 0.35)
ALL11_UPPER = (2.0, 2.0, 0.2, 1.5, 2.0, 2.0, 0.2, 1.5, 0.2, 0.8, 0.50)


def solve_state(problem: OracleProblem, theta11):
    """ParameterSensitivity::solveState (rans-uq/src/ParameterSensitivity.cpp:35-59):
    all11 bounds, a cold solve without the observation monitor. Returns
    (status, state or None)."""
    t = np.asarray(theta11, float)
    if not all(lo <= v <= hi for v, lo, hi in zip(t, ALL11_LOWER, ALL11_UPPER)):
        return 3, None
    r = problem.evaluate(t, save=True, flags=OR_FLAG_NO_MONITOR)
    return r["status"], (r["state"] if r["status"] == 0 else None)


def eta_jacobian_warm_fd(problem: OracleProblem, state, theta11, h_rel=5e-4, h_floor=1e-7,
                         warm_max_iter=0, warm_tol=0.0, absolute_steps=None):
    """ParameterSensitivity::etaJacobianWarmFD (rans-uq/src/ParameterSensitivity.cpp:729-840)
    on the oracle: central differences of re-converged solves started from the
    base state with the boundary conditions re-applied, no observation monitor
    and the per-coefficient tolerance min(tol, 0.01 h / |theta|)."""
    t0 = np.asarray(theta11, float)
    s = problem.structs[2]
    n_obs = problem.prepared.n_obs
    out = {"d_obs_d_theta": np.full((n_obs, 11), np.nan), "log_lik_gradient": np.full(11, np.nan),
           "iterations": np.zeros(11, int), "worst_measure": np.zeros(11),
           "converged": np.zeros(11, int), "steps": np.zeros(11), "tolerances": np.zeros(11),
           "positive_status": np.full(11, 4), "negative_status": np.full(11, 4),
           "n_residual_evals": 0}

    def step(j):
        if absolute_steps is not None:
            return float(absolute_steps[j])
        hr = min(h_rel, 1e-5) if j in (8, 9) else h_rel
        return max(hr * abs(t0[j]), h_floor)

    def warm_eta(tp, tol):
        if not all(np.isfinite(v) and v > 0 for v in tp):
            return 3, 0, math.inf, math.nan, None
        r = problem.evaluate(tp, warm=state, flags=OR_FLAG_NO_MONITOR | OR_FLAG_WARM_BCS,
                             tol=tol, max_iter=warm_max_iter)
        out["n_residual_evals"] += r["iters"]
        worst = r["worst"] if np.isfinite(r["worst"]) else math.inf
        if r["status"] != 0:
            return r["status"], r["iters"], worst, math.nan, None
        return 0, r["iters"], worst, r["loglik"], np.array(r["predictions"])

    for j in range(11):
        h = step(j)
        tp, tm = t0.copy(), t0.copy()
        tp[j] += h
        tm[j] -= h
        h = 0.5 * (tp[j] - tm[j])
        out["steps"][j] = h
        rel = h / max(abs(t0[j]), h_floor)
        tol = warm_tol if warm_tol > 0 else min(s.convergence_tol, 0.01 * rel)
        out["tolerances"][j] = tol
        sp, ip, wp, lp, ep = warm_eta(tp, tol)
        sm, im, wm, lm, em = warm_eta(tm, tol)
        out["positive_status"][j], out["negative_status"][j] = sp, sm
        out["iterations"][j] = ip + im
        out["worst_measure"][j] = max(wp, wm)
        ok = sp == 0 and sm == 0
        out["converged"][j] = int(ok)
        if ok:
            out["d_obs_d_theta"][:, j] = (ep - em) / (2.0 * h)
            out["log_lik_gradient"][j] = (lp - lm) / (2.0 * h)
    return out
