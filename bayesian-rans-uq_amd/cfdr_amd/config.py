"""Problem definition: settings, SST coefficients, parameter sets, boundary
conditions and the observation operator with its prepared stencils.

Names, defaults and validation follow the reference's C++ types and their
Python bindings:
  SolverSettings          core/include/FlowFields.hpp:70-241
  SSTCoefficients         core/include/SSTModel.hpp:12-61
  InferenceParameterSet   core/include/InferenceParameters.hpp:18-188
  FlowBoundaryConditions  core/include/BoundaryCondition.hpp:105-167
  ObservationOperator     core/include/ObservationOperator.hpp:65-777
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, fields

import numpy as np

from .mesh import Mesh

# --------------------------------------------------------------------------
# enums shared with the C-ABI (include/rans_amd.h)
# --------------------------------------------------------------------------
BC_DIRICHLET, BC_NEUMANN, BC_INLET, BC_OUTLET, BC_WALL_NOSLIP, BC_WALL_KOMEGA, \
    BC_WALL_MOVING, BC_SYMMETRY, BC_CYCLIC = range(9)
SOLVERS = {"PCG": 0, "BiCGSTAB": 1, "AMG": 2}
SCHEMES = {"upwind": 0, "linearUpwind": 1, "limitedLinearUpwind": 2}
OBS_VELOCITY, OBS_PRESSURE, OBS_SKIN_FRICTION = 0, 1, 2
OBS_DRAG, OBS_SEPARATION, OBS_REATTACHMENT = 3, 4, 5
PATCH_TYPES_ORDER = {"wall": 0, "inlet": 1, "outlet": 2, "symmetry": 3}

CONVERGED, UNCONVERGED, DIVERGED, INVALID_PARAMETERS, UNKNOWN = range(5)
STATUS_NAMES = ["Converged", "Unconverged", "DivergenceDetected", "InvalidParameters", "Unknown"]


@dataclass
class Vec3:
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0


# --------------------------------------------------------------------------
@dataclass
class SolverSettings:
    """SolverSettings (core/include/FlowFields.hpp:70-241), defaults kept."""

    max_iterations: int = 500
    convergence_tol: float = 1e-5
    divergence_limit: float = 1e6
    convergence_window: int = 25
    monitor_interval: int = 10
    alpha_u: float = 0.7
    alpha_p: float = 0.3
    alpha_k: float = 0.5
    alpha_omega: float = 0.5
    inner_iterations: int = 200
    inner_tolerance: float = 1e-3
    turb_start_iter: int = 5
    turb_update_interval: int = 1
    nut_floor_iters: int = 500
    momentum_convection: str = "limitedLinearUpwind"
    turbulence_convection: str = "limitedLinearUpwind"
    body_force: Vec3 = field(default_factory=Vec3)
    k_min: float = 1e-10
    k_max: float = 1e10
    omega_min: float = 1e-6
    pressure_solver: str = "AMG"
    momentum_solver: str = "BiCGSTAB"
    turbulence_solver: str = "BiCGSTAB"
    sst_variant: int = 0
    verbose: bool = False
    # reference options this path keeps at their defaults
    legacy_convergence: bool = False
    legacy_rhie_chow_gate: bool = False
    legacy_omega_wall_pinning: bool = False
    use_wall_functions: bool = False

    def validate(self) -> None:
        """SolverSettings::validate (FlowFields.hpp:203-240)."""
        if self.sst_variant not in (0, 1, 2):
            raise ValueError("unknown SST variant: expected Full (0), NoLimiter (1), or KOmega (2)")
        if (self.max_iterations < 0 or self.inner_iterations <= 0 or self.convergence_window < 2
                or self.monitor_interval <= 0 or self.turb_update_interval <= 0
                or self.turb_start_iter < 0 or self.nut_floor_iters < 0
                or self.turb_start_iter > 2**31 - 1 - self.nut_floor_iters):
            raise ValueError("invalid SIMPLE iteration count or cadence")
        if not math.isfinite(self.divergence_limit) or self.divergence_limit < 0:
            raise ValueError("SIMPLE divergence limit must be finite and nonnegative")
        if not math.isfinite(self.convergence_tol) or self.convergence_tol < 0:
            raise ValueError("SIMPLE convergence tolerance must be finite and nonnegative")
        for v in (self.inner_tolerance, self.alpha_u, self.alpha_k, self.alpha_omega):
            if not math.isfinite(v) or v <= 0:
                raise ValueError("SIMPLE tolerances and relaxation denominators must be finite and positive")
        if not math.isfinite(self.alpha_p) or self.alpha_p < 0:
            raise ValueError("SIMPLE pressure/source relaxation must be finite and nonnegative")
        b = self.body_force
        if not all(math.isfinite(v) for v in (b.x, b.y, b.z)):
            raise ValueError("SIMPLE body force must be finite")
        for v in (self.k_min, self.k_max, self.omega_min):
            if not math.isfinite(v) or v < 0:
                raise ValueError("SIMPLE turbulence limits must be finite and nonnegative")
        if self.k_max < self.k_min:
            raise ValueError("SIMPLE kMax must be at least kMin")
        for s in (self.momentum_convection, self.turbulence_convection):
            if s not in SCHEMES:
                raise ValueError("unknown SIMPLE convection scheme: " + s)
        for s in (self.pressure_solver, self.momentum_solver, self.turbulence_solver):
            if s not in SOLVERS:
                raise ValueError(f"unknown linear solver: '{s}'")
        if self.momentum_solver == "AMG" or self.turbulence_solver == "AMG":
            raise ValueError("AMG is supported for the pressure correction only")
        if (self.legacy_convergence or self.legacy_rhie_chow_gate or self.legacy_omega_wall_pinning
                or self.use_wall_functions):
            raise ValueError("legacy reproduction modes and wall functions are not on this path")
        if b.z != 0.0:
            raise ValueError("planar solver: body force must have zero z component")


# --------------------------------------------------------------------------
COEFF_NAMES = ["sigma_k1", "sigma_w1", "beta1", "alpha1", "sigma_k2", "sigma_w2", "beta2",
               "alpha2", "betaStar", "a1", "kappa"]
COEFF_DEFAULTS = [0.85, 0.5, 0.075, 5.0 / 9.0, 1.0, 0.856, 0.0828, 0.44, 0.09, 0.31, 0.41]
LOWER_BOUNDS = [0.1, 0.1, 0.01, 0.1, 0.1, 0.1, 0.01, 0.1, 0.01, 0.1, 0.35]
UPPER_BOUNDS = [2.0, 2.0, 0.2, 1.5, 2.0, 2.0, 0.2, 1.5, 0.2, 0.8, 0.50]


def validate_coefficients(c) -> None:
    """SSTCoefficients::validate (SSTModel.hpp:29-45)."""
    for name, v in zip(COEFF_NAMES, c):
        if not math.isfinite(v) or v < 0:
            raise ValueError(f"SST {name} must be finite and nonnegative")
    if c[2] == 0 or c[8] == 0:
        raise ValueError("SST beta1 and betaStar must be positive")


@dataclass
class InferenceParameterSet:
    """InferenceParameterSet (core/include/InferenceParameters.hpp:18-188)."""

    name: str
    active_indices: list
    base: list = field(default_factory=lambda: list(COEFF_DEFAULTS))

    def n_active(self) -> int:
        return len(self.active_indices)

    def validate(self) -> None:
        if not self.active_indices:
            raise ValueError("inference parameters require at least one active index")
        seen = set()
        for i in self.active_indices:
            if i < 0 or i >= 11:
                raise ValueError("inference parameter index must be in [0,10]")
            if i in seen:
                raise ValueError("inference parameter indices must be unique")
            seen.add(i)

    def unpack(self, theta) -> list:
        self.validate()
        if len(theta) != self.n_active():
            raise ValueError(f"theta size mismatch: expected {self.n_active()}, got {len(theta)}")
        full = list(self.base)
        for i, idx in enumerate(self.active_indices):
            full[idx] = float(theta[i])
        return full

    def lower_bounds(self):
        return [LOWER_BOUNDS[i] for i in self.active_indices]

    def upper_bounds(self):
        return [UPPER_BOUNDS[i] for i in self.active_indices]

    def active_names(self):
        return [COEFF_NAMES[i] for i in self.active_indices]

    def in_bounds(self, theta) -> bool:
        self.validate()
        if len(theta) != self.n_active():
            return False
        for v, lo, hi in zip(theta, self.lower_bounds(), self.upper_bounds()):
            if v < lo or v > hi:
                return False
            if math.isnan(v) or math.isinf(v):
                return False
        return True

    @staticmethod
    def a1_betaStar():
        return InferenceParameterSet("a1_betaStar", [9, 8])

    @staticmethod
    def inlet_turb():
        return InferenceParameterSet("inletTurb", [2, 1])

    @staticmethod
    def near_wall4():
        return InferenceParameterSet("nearWall4", [9, 8, 2, 0])

    @staticmethod
    def all11():
        return InferenceParameterSet("all11", list(range(11)))

    @staticmethod
    def live10():
        return InferenceParameterSet("live10", list(range(10)))

    @staticmethod
    def from_indices(name, idx):
        if not idx:
            raise ValueError("fromIndices: need at least one active index")
        for i, a in enumerate(idx):
            if a < 0 or a > 10:
                raise ValueError("fromIndices: index out of range [0,10]")
            if a in idx[i + 1:]:
                raise ValueError("fromIndices: duplicate index")
        return InferenceParameterSet(name, list(idx))


# --------------------------------------------------------------------------
@dataclass
class PatchBC:
    type: int = BC_NEUMANN
    value: float = 0.0
    vec_value: Vec3 = field(default_factory=Vec3)


@dataclass
class FlowBoundaryConditions:
    """FlowBoundaryConditions (BoundaryCondition.hpp:105-167): one PatchBC per
    mesh patch for velocity, pressure, k and omega (uniform values)."""

    velocity: list
    pressure: list
    k: list
    omega: list

    @staticmethod
    def channel_defaults(mesh: Mesh, u_in: float, k_in: float, om_in: float):
        vel, pre, kk, om = [], [], [], []
        for p in mesh.patches:
            if p.type == "wall":
                vel.append(PatchBC(BC_WALL_NOSLIP))
                pre.append(PatchBC(BC_NEUMANN))
                kk.append(PatchBC(BC_WALL_KOMEGA))
                om.append(PatchBC(BC_WALL_KOMEGA))
            elif p.type == "inlet":
                vel.append(PatchBC(BC_INLET, 0.0, Vec3(u_in, 0, 0)))
                pre.append(PatchBC(BC_NEUMANN))
                kk.append(PatchBC(BC_DIRICHLET, k_in))
                om.append(PatchBC(BC_DIRICHLET, om_in))
            elif p.type == "outlet":
                vel.append(PatchBC(BC_NEUMANN))
                pre.append(PatchBC(BC_OUTLET, 0.0))
                kk.append(PatchBC(BC_NEUMANN))
                om.append(PatchBC(BC_NEUMANN))
            elif p.type == "symmetry":
                vel.append(PatchBC(BC_SYMMETRY))
                pre.append(PatchBC(BC_NEUMANN))
                kk.append(PatchBC(BC_NEUMANN))
                om.append(PatchBC(BC_NEUMANN))
            else:
                vel.append(PatchBC())
                pre.append(PatchBC())
                kk.append(PatchBC())
                om.append(PatchBC())
        return FlowBoundaryConditions(vel, pre, kk, om)

    @staticmethod
    def couette_defaults(mesh: Mesh, u_wall: float, k_in: float, om_in: float):
        """couetteDefaults (BoundaryCondition.hpp:204-222): the channel
        defaults at half the wall speed, the patch named 'top' a wall moving
        at (u_wall, 0, 0) (k and omega keep their wall treatment), and the
        patch named 'inlet' zero-gradient in every field."""
        bc = FlowBoundaryConditions.channel_defaults(mesh, 0.5 * u_wall, k_in, om_in)
        for i, p in enumerate(mesh.patches):
            if p.name == "top":
                bc.velocity[i] = PatchBC(BC_WALL_MOVING, 0.0, Vec3(u_wall, 0, 0))
            elif p.name == "inlet":
                bc.velocity[i] = PatchBC(BC_NEUMANN)
                bc.pressure[i] = PatchBC(BC_NEUMANN)
                bc.k[i] = PatchBC(BC_NEUMANN)
                bc.omega[i] = PatchBC(BC_NEUMANN)
        return bc

    @staticmethod
    def flat_plate_defaults(mesh: Mesh, u_inf: float, k_in: float, om_in: float):
        """flatPlateDefaults (BoundaryCondition.hpp:177-192): the channel
        defaults, with the patch named 'top' a free stream: Dirichlet velocity
        (u_inf, 0, 0), k and omega at their inlet values, zero-gradient
        pressure."""
        bc = FlowBoundaryConditions.channel_defaults(mesh, u_inf, k_in, om_in)
        for i, p in enumerate(mesh.patches):
            if p.name == "top":
                bc.velocity[i] = PatchBC(BC_DIRICHLET, 0.0, Vec3(u_inf, 0, 0))
                bc.pressure[i] = PatchBC(BC_NEUMANN)
                bc.k[i] = PatchBC(BC_DIRICHLET, k_in)
                bc.omega[i] = PatchBC(BC_DIRICHLET, om_in)
        return bc

    def validate(self, mesh: Mesh) -> None:
        """validatePatchBCs + simple_admission::turbulenceBoundary
        (BoundaryCondition.hpp:45-103, SIMPLEValidation.hpp:14-31)."""
        n = mesh.n_patches()
        for lst in (self.velocity, self.pressure, self.k, self.omega):
            if len(lst) != n:
                raise ValueError("boundary conditions require one entry per mesh patch")
            for bc in lst:
                if bc.type not in range(9):
                    raise ValueError("unknown boundary condition type")
        for bc in self.velocity:
            if bc.type in (BC_DIRICHLET, BC_INLET, BC_WALL_MOVING):
                v = bc.vec_value
                if not all(math.isfinite(t) for t in (v.x, v.y, v.z)):
                    raise ValueError("active vector boundary values must be finite")
                if v.z != 0.0:
                    raise ValueError("planar solver: boundary velocity must have zero z component")
        for lst in (self.pressure, self.k, self.omega):
            for bc in lst:
                if bc.type in (BC_DIRICHLET, BC_OUTLET) and not math.isfinite(bc.value):
                    raise ValueError("active scalar boundary values must be finite")
        for lst in (self.k, self.omega):
            for bc in lst:
                if bc.type == BC_DIRICHLET and bc.value < 0:
                    raise ValueError("SIMPLE prescribed turbulence must be nonnegative")

    def arrays(self):
        """Per-patch type codes and values as flat arrays for the C-ABI."""
        ut = np.asarray([b.type for b in self.velocity], np.int32)
        uv = np.asarray([(b.vec_value.x, b.vec_value.y) for b in self.velocity], float).reshape(-1)
        pt = np.asarray([b.type for b in self.pressure], np.int32)
        pv = np.asarray([b.value for b in self.pressure], float)
        kt = np.asarray([b.type for b in self.k], np.int32)
        kv = np.asarray([b.value for b in self.k], float)
        wt = np.asarray([b.type for b in self.omega], np.int32)
        wv = np.asarray([b.value for b in self.omega], float)
        return ut, uv, pt, pv, kt, kv, wt, wv


# --------------------------------------------------------------------------
@dataclass
class Observable:
    type: int
    observed: float
    sigma_obs: float
    sigma_model: float = 0.0
    location: tuple = (0.0, 0.0, 0.0)
    component: int = 0
    patch_name: str = ""
    ref_velocity: float = 1.0
    ref_pressure: float = 0.0
    ref_density: float = 1.0
    reference_area: float = 1.0

    def total_variance(self) -> float:
        return self.sigma_obs * self.sigma_obs + self.sigma_model * self.sigma_model

    def validate(self) -> None:
        tv = self.total_variance()
        if (not math.isfinite(self.observed) or not math.isfinite(self.sigma_obs)
                or self.sigma_obs < 0 or not math.isfinite(self.sigma_model)
                or self.sigma_model < 0 or not math.isfinite(tv) or tv <= 0):
            raise ValueError("Gaussian measurements require finite values, nonnegative "
                             "uncertainties and finite positive total variance")


class ObservationOperator:
    """ObservationOperator (ObservationOperator.hpp:65-150): velocity probes,
    pressure taps, skin friction, drag, separation points and reattachment
    lengths."""

    def __init__(self):
        self.observables: list[Observable] = []

    def n_obs(self) -> int:
        return len(self.observables)

    @staticmethod
    def _loc(loc):
        if isinstance(loc, Vec3):
            loc = (loc.x, loc.y, loc.z)
        loc = tuple(float(v) for v in loc)
        if len(loc) == 2:
            loc = loc + (0.0,)
        if not all(math.isfinite(v) for v in loc):
            raise ValueError("observation location must be finite")
        return loc

    def add_velocity_profile(self, loc, comp: int, u_obs: float, sigma: float,
                             sigma_model: float = 0.0):
        loc = self._loc(loc)
        if comp < 0 or comp > 2:
            raise ValueError("velocity observation component must be in [0,2]")
        ob = Observable(OBS_VELOCITY, u_obs, sigma, sigma_model, loc, comp)
        ob.validate()
        self.observables.append(ob)

    def add_pressure_tap(self, loc, cp_obs: float, sigma: float, ref_vel: float,
                         sigma_model: float = 0.0, ref_pressure: float = 0.0,
                         ref_density: float = 1.0):
        loc = self._loc(loc)
        if not math.isfinite(ref_vel) or ref_vel <= 0:
            raise ValueError("pressure reference speed must be finite and positive")
        if not math.isfinite(ref_density) or ref_density <= 0:
            raise ValueError("pressure reference density must be finite and positive")
        ob = Observable(OBS_PRESSURE, cp_obs, sigma, sigma_model, loc, 0, "", ref_vel,
                        ref_pressure, ref_density)
        ob.validate()
        self.observables.append(ob)

    def add_skin_friction(self, patch: str, loc, cf_obs: float, sigma: float, ref_vel: float,
                          sigma_model: float = 0.0):
        loc = self._loc(loc)
        if not math.isfinite(ref_vel) or ref_vel <= 0:
            raise ValueError("friction reference speed must be finite and positive")
        ob = Observable(OBS_SKIN_FRICTION, cf_obs, sigma, sigma_model, loc, 0, patch, ref_vel)
        ob.validate()
        self.observables.append(ob)

    def add_drag(self, patch: str, cd_obs: float, sigma: float, ref_area: float, ref_vel: float,
                 sigma_model: float = 0.0):
        """addDrag (ObservationOperator.hpp:189-201): the pressure and
        viscous x force on `patch` over 0.5 ref_vel^2 ref_area."""
        if not math.isfinite(ref_area) or ref_area <= 0:
            raise ValueError("drag reference area must be finite and positive")
        if not math.isfinite(ref_vel) or ref_vel <= 0:
            raise ValueError("drag reference speed must be finite and positive")
        ob = Observable(OBS_DRAG, cd_obs, sigma, sigma_model, patch_name=patch,
                        ref_velocity=ref_vel, reference_area=ref_area)
        ob.validate()
        self.observables.append(ob)

    def add_separation_point(self, patch: str, x_obs: float, sigma: float,
                             sigma_model: float = 0.0):
        """addSeparationPoint (ObservationOperator.hpp:251-259): the x where
        the streamwise wall shear on `patch` first turns non-positive."""
        ob = Observable(OBS_SEPARATION, x_obs, sigma, sigma_model, patch_name=patch)
        ob.validate()
        self.observables.append(ob)

    def add_reattachment_length(self, patch: str, x_obs: float, sigma: float,
                                sigma_model: float = 0.0):
        """addReattachmentLength (ObservationOperator.hpp:261-269): the x
        where the streamwise wall shear on `patch` first turns non-negative
        (the last face centre when it never does)."""
        ob = Observable(OBS_REATTACHMENT, x_obs, sigma, sigma_model, patch_name=patch)
        ob.validate()
        self.observables.append(ob)

    def prepare(self, mesh: Mesh) -> "PreparedObservation":
        return PreparedObservation(mesh, self)

    def log_likelihood(self, predictions) -> float:
        """ObservationOperator::logLikelihood (ObservationOperator.hpp:592-607)."""
        if len(predictions) != len(self.observables):
            raise ValueError("Likelihood prediction count does not match observations")
        ll = 0.0
        for ob, pred in zip(self.observables, predictions):
            ob.validate()
            if not math.isfinite(pred):
                raise ValueError("Likelihood predictions must be finite")
            diff = pred - ob.observed
            ll += -0.5 * diff * diff / ob.total_variance()
            if not math.isfinite(ll):
                raise OverflowError("Gaussian log-likelihood exceeds finite range")
        return ll


# ---- geometric search helpers (ObservationOperator.hpp:272-322, 384-419) ---
def _nearest_cell(mesh: Mesh, loc) -> int:
    c = mesh.cell_center
    dx, dy, dz = c[:, 0] - loc[0], c[:, 1] - loc[1], c[:, 2] - loc[2]
    d = dx * dx + dy * dy + dz * dz
    best, best_d = 0, 1e30
    for ci in range(mesh.n_cells):  # first minimum, strict comparison
        if d[ci] < best_d:
            best_d, best = d[ci], ci
    return best


def _containing_cell(mesh: Mesh, loc):
    ci = _nearest_cell(mesh, loc)
    lx, ly, lz = loc
    for _ in range(2 * mesh.n_cells + 10):
        worst, cross = 0.0, -1
        for fi in mesh.cell_face_idx[mesh.cell_face_ptr[ci]:mesh.cell_face_ptr[ci + 1]]:
            n = mesh.face_normal[fi]
            if abs(n[2]) > 0.5:
                continue
            own = mesh.owner[fi] == ci
            nx, ny = (n[0], n[1]) if own else (n[0] * -1.0, n[1] * -1.0)
            cf = mesh.face_center[fi] if own else mesh.face_center[fi] - mesh.face_shift[fi]
            out = (lx - cf[0]) * nx + (ly - cf[1]) * ny
            if out > worst + 1e-12 * math.sqrt(mesh.face_area[fi]):
                worst, cross = out, int(fi)
        if cross < 0:
            return ci, (lx, ly, lz)
        if mesh.neighbor[cross] < 0:
            return ci, (lx, ly, lz)
        own = mesh.owner[cross] == ci
        s = mesh.face_shift[cross]
        if own:
            lx, ly, lz = lx - s[0], ly - s[1], lz - s[2]
            ci = int(mesh.neighbor[cross])
        else:
            lx, ly, lz = lx + s[0], ly + s[1], lz + s[2]
            ci = int(mesh.owner[cross])
    return _nearest_cell(mesh, loc), tuple(loc)


def _nearest_wall_face(mesh: Mesh, patch: str, loc) -> int:
    faces = mesh.patches[mesh.patch_by_name(patch)].faces
    best = int(faces[0]) if len(faces) else -1
    best_d = 1e30
    for fi in faces:
        c = mesh.face_center[fi]
        d = (c[0] - loc[0]) * (c[0] - loc[0]) + (c[1] - loc[1]) * (c[1] - loc[1]) + \
            (c[2] - loc[2]) * (c[2] - loc[2])
        if d < best_d:
            best_d, best = d, int(fi)
    return best


def _friction_stencil(mesh: Mesh, ob: Observable):
    """ObservationOperator::skinFrictionStencil (ObservationOperator.hpp:384-412)."""
    fa = _nearest_wall_face(mesh, ob.patch_name, ob.location)
    if fa < 0:
        return -1, -1, 0.0
    faces = mesh.patches[mesh.patch_by_name(ob.patch_name)].faces
    ca = mesh.face_center[fa]
    dl = (ob.location[0] - ca[0], ob.location[1] - ca[1], ob.location[2] - ca[2])
    best, fb = 1e300, -1
    for fi in faces:
        if fi == fa:
            continue
        c = mesh.face_center[fi]
        e = (c[0] - ca[0], c[1] - ca[1], c[2] - ca[2])
        if e[0] * dl[0] + e[1] * dl[1] + e[2] * dl[2] <= 0.0:
            continue
        n2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2]
        if n2 < best:
            best, fb = n2, int(fi)
    t = 0.0
    if fb >= 0:
        c = mesh.face_center[fb]
        e = (c[0] - ca[0], c[1] - ca[1], c[2] - ca[2])
        n2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2]
        t = min(max((dl[0] * e[0] + dl[1] * e[1] + dl[2] * e[2]) / n2, 0.0), 1.0)
    return fa, fb, t


class PreparedObservation:
    """PreparedObservation (ObservationOperator.hpp:611-669): precomputed
    probe stencils (containing cell, offset, least-squares neighbours and
    their 1/|d|^2 weights) and skin-friction face pairs, flattened for the
    device gather."""

    def __init__(self, mesh: Mesh, op: ObservationOperator):
        n = op.n_obs()
        self.n_obs = n
        self.type = np.asarray([o.type for o in op.observables], np.int32)
        self.component = np.asarray([o.component for o in op.observables], np.int32)
        self.observed = np.asarray([o.observed for o in op.observables], float)
        self.variance = np.asarray([o.total_variance() for o in op.observables], float)
        self.ref_velocity = np.asarray([o.ref_velocity for o in op.observables], float)
        self.ref_pressure = np.asarray([o.ref_pressure for o in op.observables], float)
        self.ref_density = np.asarray([o.ref_density for o in op.observables], float)
        self.cell = np.full(n, -1, np.int32)
        self.offx, self.offy = np.zeros(n), np.zeros(n)
        self.axx, self.axy, self.ayy, self.det = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        self.face_a = np.full(n, -1, np.int32)
        self.face_b = np.full(n, -1, np.int32)
        self.t = np.zeros(n)
        nb_ptr = [0]
        nb_cell, nb_face, nb_dx, nb_dy, nb_w = [], [], [], [], []
        for i, ob in enumerate(op.observables):
            if ob.type in (OBS_VELOCITY, OBS_PRESSURE):
                if mesh.n_cells == 0:
                    raise ValueError("cannot prepare a probe on an empty mesh")
                ci, pt = _containing_cell(mesh, ob.location)
                c = mesh.cell_center[ci]
                self.cell[i] = ci
                self.offx[i], self.offy[i] = pt[0] - c[0], pt[1] - c[1]
                axx = axy = ayy = 0.0
                for fi in mesh.cell_face_idx[mesh.cell_face_ptr[ci]:mesh.cell_face_ptr[ci + 1]]:
                    if abs(mesh.face_normal[fi][2]) > .5:
                        continue
                    if mesh.neighbor[fi] < 0:
                        fc = mesh.face_center[fi]
                        dx, dy = fc[0] - c[0], fc[1] - c[1]
                        nb = -1
                    else:
                        own = mesh.owner[fi] == ci
                        nb = int(mesh.neighbor[fi] if own else mesh.owner[fi])
                        cn, s = mesh.cell_center[nb], mesh.face_shift[fi]
                        if own:
                            dx, dy = (cn[0] + s[0]) - c[0], (cn[1] + s[1]) - c[1]
                        else:
                            dx, dy = (cn[0] - s[0]) - c[0], (cn[1] - s[1]) - c[1]
                    w = 1.0 / max(dx * dx + dy * dy, 1e-300)
                    axx += w * dx * dx
                    axy += w * dx * dy
                    ayy += w * dy * dy
                    nb_cell.append(nb)
                    nb_face.append(int(fi))
                    nb_dx.append(dx)
                    nb_dy.append(dy)
                    nb_w.append(w)
                self.axx[i], self.axy[i], self.ayy[i] = axx, axy, ayy
                self.det[i] = axx * ayy - axy * axy
            elif ob.type == OBS_SKIN_FRICTION:
                self.face_a[i], self.face_b[i], self.t[i] = _friction_stencil(mesh, ob)
            elif ob.type in (OBS_DRAG, OBS_SEPARATION, OBS_REATTACHMENT):
                # the patch's faces in patch order (ObservationOperator.hpp:665)
                self.offx[i] = ob.reference_area
                pat = [p for p in mesh.patches if p.name == ob.patch_name]
                if not pat:
                    raise ValueError(f"unknown patch '{ob.patch_name}'")
                for fi in pat[0].faces:
                    nb_cell.append(-1)
                    nb_face.append(int(fi))
                    nb_dx.append(float(mesh.face_center[fi][0]))
                    nb_dy.append(0.0)
                    nb_w.append(0.0)
            nb_ptr.append(len(nb_cell))
        self.nb_ptr = np.asarray(nb_ptr, np.int32)
        self.nb_cell = np.asarray(nb_cell, np.int32)
        self.nb_face = np.asarray(nb_face, np.int32)
        self.nb_dx = np.asarray(nb_dx, float)
        self.nb_dy = np.asarray(nb_dy, float)
        self.nb_w = np.asarray(nb_w, float)
