"""The forward-solve workloads as concrete synthetic problems.

* ``channel_problem``: fully developed turbulent channel at Re_tau = 182.088
  in u_tau = 1 units (nu = 1/Re_tau, body force 1), one streamwise column of
  ``ny`` wall-normal cells with geometric growth from both walls, SST
  coefficients at their defaults, Jacobi-PCG pressure and BiCGSTAB transport,
  velocity probes at y+ uniformly spaced in [1, 182]. Settings follow the
  reference's calibration driver (docs/figures/calibration_coverage.py:95-112).
* ``channel_draws``: seeded bivariate Student-t (df = 4) draws of (a1, beta*).
* ``hill_problem``: periodic hill (ERCOFTAC polynomial, L = 9h, top 3.035h,
  Re_h = 5600) with the reference's hills settings
  (tests/tools/record_goldens.cpp:558-591).
* ``hill_draws``: uniform draws of (a1, beta*, sigma_k1, sigma_w1) within
  +/-10% of the SST defaults.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .config import (FlowBoundaryConditions, InferenceParameterSet, ObservationOperator,
                     SolverSettings, Vec3)
from .mesh import Mesh, hill_bottom, make_hill_mesh

RE_TAU = 182.088
CHANNEL_SEED = 20261010
HILL_SEED = 5600


@dataclass
class Problem:
    mesh: Mesh
    params: InferenceParameterSet
    obs: ObservationOperator
    bcs: FlowBoundaryConditions
    nu: float
    settings: SolverSettings
    u_init: Vec3
    p_init: float
    k_init: float
    omega_init: float

    def init_vector(self):
        return [self.u_init.x, self.u_init.y, self.p_init, self.k_init, self.omega_init]

    def coefficients(self, theta):
        return self.params.unpack(theta)


def channel_problem(ny: int = 128, nx: int = 1, ratio: float = 1.05, n_probes: int = 30,
                    tol: float = 1e-8, max_iterations: int = 150000,
                    pressure_solver: str = "PCG") -> Problem:
    mesh = Mesh.make_geometric_channel_2d(nx, ny, 1.0, 1.0, ratio)
    mesh.compute_wall_distance()
    bcs = FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1.0, 10.0)
    obs = ObservationOperator()
    for i in range(n_probes):
        yp = 1.0 + i * (182.0 - 1.0) / (n_probes - 1) if n_probes > 1 else 1.0
        obs.add_velocity_profile((0.5, yp / RE_TAU, 0.0), 0, 0.0, 0.1)
    s = SolverSettings(max_iterations=max_iterations, convergence_tol=tol, alpha_u=0.5,
                       body_force=Vec3(1.0, 0.0, 0.0), pressure_solver=pressure_solver)
    return Problem(mesh, InferenceParameterSet.a1_betaStar(), obs, bcs, 1.0 / RE_TAU, s,
                   Vec3(15.0, 0.0, 0.0), 0.0, 1.0, 10.0)


def channel_draws(n: int = 64, seed: int = CHANNEL_SEED) -> np.ndarray:
    """theta = mu + s * z * sqrt(df / chi2_df), df = 4, mu = (0.324, 0.0968),
    s = (0.004, 0.0015): the calibration's importance-sampling proposal
    (docs/figures/calibration_coverage.py:165-181), from numpy's PCG64."""
    rng = np.random.default_rng(seed)
    mu = np.array([0.324, 0.0968])
    scale = np.array([0.004, 0.0015])
    z = rng.standard_normal((n, 2))
    chi2 = rng.chisquare(4.0, size=(n, 1))
    return mu + scale * z * np.sqrt(4.0 / chi2)


def channel_ratio(ny: int, first: float = 0.05 / (1.05 ** 64 - 1)) -> float:
    """Geometric ratio that gives a ny-cell channel (ny/2 cells per half
    height h = 1) the first-cell height of the 128-cell, ratio-1.05 mesh."""
    half = ny // 2
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first * (mid ** half - 1.0) / (mid - 1.0) > 1.0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


SERVED_CHANNELS = (64, 128, 256, 512)
SERVED_SEED = 4242


def served_stream(n: int, rate: float, meshes=SERVED_CHANNELS, seed: int = SERVED_SEED):
    """Seeded Poisson arrivals of single-solve requests: (arrival time s,
    mesh key, theta) with the mesh drawn uniformly from `meshes` and theta
    uniform over a1 in [0.28, 0.34], beta* in [0.084, 0.096]
    (docs/figures/calibration_coverage.py:51)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    which = rng.integers(0, len(meshes), size=n)
    theta = np.column_stack([rng.uniform(0.28, 0.34, size=n), rng.uniform(0.084, 0.096, size=n)])
    return [(float(t[i]), meshes[which[i]], theta[i]) for i in range(n)]


def hill_problem(nx: int = 120, ny: int = 64, tol: float = 1e-4, max_iterations: int = 25000,
                 pressure_solver: str = "AMG") -> Problem:
    mesh = make_hill_mesh(nx, ny)
    mesh.compute_wall_distance()
    bcs = FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 10.0)
    obs = ObservationOperator()
    for x in (0.5, 2.0, 4.0, 6.0, 8.0):
        for dy in (0.1, 0.5, 1.0, 2.0):
            obs.add_velocity_profile((x, hill_bottom(x) + dy, 0.0), 0, 0.0, 0.01)
    s = SolverSettings(max_iterations=max_iterations, convergence_tol=tol, divergence_limit=1e8,
                       alpha_u=0.3, alpha_p=0.2, alpha_k=0.4, alpha_omega=0.4,
                       inner_iterations=300, inner_tolerance=1e-4, turb_start_iter=30,
                       turb_update_interval=2, body_force=Vec3(0.00906, 0.0, 0.0),
                       pressure_solver=pressure_solver)
    params = InferenceParameterSet.from_indices("hill4", [9, 8, 0, 1])
    return Problem(mesh, params, obs, bcs, 1.0 / 5600.0, s, Vec3(0.0, 0.0, 0.0), 0.0, 1e-4, 10.0)


def hill_draws(n: int = 16, seed: int = HILL_SEED) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centre = np.array([0.31, 0.09, 0.85, 0.5])
    return centre * (1.0 + rng.uniform(-0.1, 0.1, size=(n, 4)))
