"""Shared loaders for the committed golden fixtures and the problems they were
recorded on (the generator script is tests/golden/make_fixtures.py)."""
from __future__ import annotations

import json
import os

import numpy as np

from cfdr_amd import cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELD_NAMES = ["Ux", "Uy", "p", "k", "omega", "nuT", "F1", "F2", "Pk", "CDkw"]


def load(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def small_channel(max_iterations=40000):
    """The 1x32 geometric channel (ratio 1.15) every small fixture uses."""
    return cases.channel_problem(ny=32, ratio=1.15, n_probes=30, tol=1e-6,
                                 max_iterations=max_iterations)


def small_hill(pressure_solver):
    return cases.hill_problem(nx=30, ny=16, max_iterations=400, pressure_solver=pressure_solver)


# velocity probes (x, y, component, observed) and pressure taps (x, y) of the
# developing inlet/outlet channel
INOUT_PROBES = [(3.0, 0.5, 0, 1.0), (5.5, 0.1, 0, 1.0), (5.5, 0.02, 0, 1.0), (2.0, 0.3, 1, 0.0)]
INOUT_TAPS = [(1.0, 0.5), (5.0, 0.5)]


def inout_channel(momentum_scheme, turbulence_scheme):
    """The developing channel of ref_channel_inout.json: the reference's
    makeChannel2D(24, 12, 6, 1) mesh (MSH2 fixture), inlet U 1, k 1e-3,
    omega 5, nu 1e-3, tol 1e-5, PCG pressure."""
    from cfdr_amd.config import (FlowBoundaryConditions, InferenceParameterSet,
                                 ObservationOperator, SolverSettings, Vec3)
    from cfdr_amd.mesh import Mesh
    mesh = Mesh.load_binary(os.path.join(GOLDEN, "ref_channel_inout_24x12.msh"))
    mesh.compute_wall_distance()
    bcs = FlowBoundaryConditions.channel_defaults(mesh, 1.0, 0.001, 5.0)
    s = SolverSettings(max_iterations=6000, convergence_tol=1e-5, pressure_solver="PCG",
                       alpha_u=0.7, alpha_p=0.3, momentum_convection=momentum_scheme,
                       turbulence_convection=turbulence_scheme)
    obs = ObservationOperator()
    for x, y, comp, observed in INOUT_PROBES:
        obs.add_velocity_profile((x, y), comp, observed, 0.1)
    for x, y in INOUT_TAPS:
        obs.add_pressure_tap((x, y), 0.0, 0.1, 1.0, 0.0, 0.0, 1.0)
    obs.add_skin_friction("bottom", (4.0, 0.0), 0.005, 0.001, 1.0)
    return cases.Problem(mesh=mesh, bcs=bcs, settings=s, obs=obs,
                         params=InferenceParameterSet.a1_betaStar(), nu=0.001,
                         u_init=Vec3(1.0, 0.0, 0.0), p_init=0.0, k_init=0.001, omega_init=5.0)


def foam_channel(max_iterations=6000):
    """The developing channel of ref_channel_inout.json (limitedLinearUpwind,
    PCG pressure, tol 1e-5) on the OpenFOAM 2D case foam_channel_24x12: a
    graded 24x12 mesh of trapezoids, depth 0.5, its front and back faces an
    "empty" patch (ref_foam_channel.json, the reference's own solve of the
    case as its Mesh::loadFromFile reads it)."""
    from cfdr_amd.config import (FlowBoundaryConditions, InferenceParameterSet,
                                 ObservationOperator, SolverSettings, Vec3)
    from cfdr_amd.mesh import Mesh
    mesh = Mesh.load_from_file(os.path.join(GOLDEN, "foam_channel_24x12"))
    mesh.compute_wall_distance()
    bcs = FlowBoundaryConditions.channel_defaults(mesh, 1.0, 0.001, 5.0)
    s = SolverSettings(max_iterations=max_iterations, convergence_tol=1e-5,
                       pressure_solver="PCG", alpha_u=0.7, alpha_p=0.3)
    obs = ObservationOperator()
    for x, y, comp, observed in INOUT_PROBES:
        obs.add_velocity_profile((x, y), comp, observed, 0.1)
    for x, y in INOUT_TAPS:
        obs.add_pressure_tap((x, y), 0.0, 0.1, 1.0, 0.0, 0.0, 1.0)
    obs.add_skin_friction("bottom", (4.0, 0.0), 0.005, 0.001, 1.0)
    return cases.Problem(mesh=mesh, bcs=bcs, settings=s, obs=obs,
                         params=InferenceParameterSet.a1_betaStar(), nu=0.001,
                         u_init=Vec3(1.0, 0.0, 0.0), p_init=0.0, k_init=0.001, omega_init=5.0)


def couette_problem(pressure_solver="PCG"):
    """Moving-wall Couette flow of ref_couette.json: the reference's graded
    makeChannel2D(4, 96, 1.5, 2) mesh (MSH2 fixture), top wall at Uwall 2,
    couetteDefaults, skin friction on 'bottom' and 18 velocity stations at
    x = 0.75, tol 1e-4, alpha_u 0.7, alpha_p 0.5 (the reference's Couette
    regression, tests/tools/record_goldens.cpp:277-301)."""
    from cfdr_amd.config import (FlowBoundaryConditions, InferenceParameterSet,
                                 ObservationOperator, SolverSettings, Vec3)
    from cfdr_amd.mesh import Mesh
    d = load("ref_couette.json")
    mesh = Mesh.load_binary(os.path.join(GOLDEN, "ref_couette_4x96.msh"))
    mesh.compute_wall_distance()
    bcs = FlowBoundaryConditions.couette_defaults(mesh, d["u_wall"], d["k_in"], d["omega_in"])
    s = SolverSettings(max_iterations=12000, convergence_tol=1e-4, alpha_u=0.7, alpha_p=0.5,
                       pressure_solver=pressure_solver)
    obs = ObservationOperator()
    obs.add_skin_friction("bottom", (0.75, 0.0), 6e-3, 1e-3, d["u_bulk"])
    for y in d["stations"]:
        obs.add_velocity_profile((0.75, y), 0, 0.0, 0.01)
    return cases.Problem(mesh=mesh, bcs=bcs, settings=s, obs=obs,
                         params=InferenceParameterSet.a1_betaStar(), nu=d["nu"],
                         u_init=Vec3(d["u_bulk"], 0.0, 0.0), p_init=0.0, k_init=d["k_in"],
                         omega_init=d["omega_in"])


def plate_problem():
    """Flat plate of ref_plate.json: the reference's makePlate2D(24, 16, 1,
    0.1, 1e5, 1) mesh (MSH2 fixture), flatPlateDefaults (free-stream top),
    the plate-validation settings of rans-uq/apps/Main.cpp:137-176 with PCG
    pressure, skin friction along the plate and a velocity profile at x 0.5."""
    from cfdr_amd.config import (FlowBoundaryConditions, InferenceParameterSet,
                                 ObservationOperator, SolverSettings, Vec3)
    from cfdr_amd.mesh import Mesh
    d = load("ref_plate.json")
    mesh = Mesh.load_binary(os.path.join(GOLDEN, "ref_plate_24x16.msh"))
    mesh.compute_wall_distance()
    bcs = FlowBoundaryConditions.flat_plate_defaults(mesh, d["u_inf"], d["k_in"], d["omega_in"])
    s = SolverSettings(max_iterations=d["max_iterations"], convergence_tol=1e-4, alpha_u=0.7,
                       alpha_p=0.5, turb_start_iter=0, k_max=0.1, pressure_solver="PCG")
    obs = ObservationOperator()
    for x in d["cf_x"]:
        obs.add_skin_friction("bottom", (x, 0.0), 0.004, 0.001, d["u_inf"])
    for y in d["stations"]:
        obs.add_velocity_profile((0.5, y), 0, 0.9, 0.01)
    return cases.Problem(mesh=mesh, bcs=bcs, settings=s, obs=obs,
                         params=InferenceParameterSet.a1_betaStar(), nu=d["nu"],
                         u_init=Vec3(d["u_inf"], 0.0, 0.0), p_init=0.0, k_init=d["k_in"],
                         omega_init=d["omega_in"])


def bfs_problem():
    """Backward-facing step of ref_bfs.json: the reference's
    makeBackwardFacingStep2D(20, 24, 12, 9, 10, 22, 1, 6, 5100, 1) mesh with a
    symmetry top (MSH2 fixture), bfsDefaults (the channel defaults), the
    settings of the reference's Le-Moin-Kim regression with PCG pressure and
    600 capped iterations, pressure taps and skin friction along the lower
    wall behind the step, and a velocity profile at x 4."""
    from cfdr_amd.config import (FlowBoundaryConditions, InferenceParameterSet,
                                 ObservationOperator, SolverSettings, Vec3)
    from cfdr_amd.mesh import Mesh
    d = load("ref_bfs.json")
    mesh = Mesh.load_binary(os.path.join(GOLDEN, "ref_bfs_20x24.msh"))
    mesh.compute_wall_distance()
    bcs = FlowBoundaryConditions.channel_defaults(mesh, d["u_bulk"], d["k_in"], d["omega_in"])
    s = SolverSettings(max_iterations=d["max_iterations"], convergence_tol=1e-4,
                       divergence_limit=1e8, alpha_u=0.3, alpha_p=0.2, alpha_k=0.4,
                       alpha_omega=0.4, inner_iterations=300, inner_tolerance=1e-4,
                       turb_start_iter=30, turb_update_interval=2, pressure_solver="PCG")
    obs = ObservationOperator()
    for x in d["taps"]:
        obs.add_pressure_tap((x, 0.0), 0.0, 0.01, d["u_bulk"], 0.0, 0.0, 1.0)
    for x in d["taps"]:
        obs.add_skin_friction("bottom_wall_down", (x, 0.0), 0.0, 1e-3, d["u_bulk"])
    for y in d["profile_y"]:
        obs.add_velocity_profile((4.0, y), 0, 0.0, 0.01)
    return cases.Problem(mesh=mesh, bcs=bcs, settings=s, obs=obs,
                         params=InferenceParameterSet.a1_betaStar(), nu=d["nu"],
                         u_init=Vec3(d["u_bulk"], 0.0, 0.0), p_init=0.0, k_init=d["k_in"],
                         omega_init=d["omega_in"])


def crossings_problems():
    """The two problems of ref_wall_crossings.json: the periodic cosine bump
    (separation point, reattachment length and a velocity probe) and the
    backward-facing step of bfs_problem() observed by a reattachment length,
    two separation points and a velocity probe."""
    from cfdr_amd.config import (FlowBoundaryConditions, InferenceParameterSet,
                                 ObservationOperator, SolverSettings, Vec3)
    from cfdr_amd.mesh import Mesh
    d = load("ref_wall_crossings.json")
    c = d["cosine_bump"]
    mesh = Mesh.make_curved_channel_periodic_2d(c["x_nodes"], c["y_bottom"], c["y_top"], c["ny"],
                                                0.0, 1.0)
    mesh.compute_wall_distance()
    s = SolverSettings(max_iterations=1500, convergence_tol=1e-4, alpha_u=0.4, alpha_p=0.25,
                       body_force=Vec3(0.004, 0.0, 0.0), pressure_solver="PCG")
    obs = ObservationOperator()
    obs.add_separation_point("bottom_wall", 0.0, 0.1)
    obs.add_reattachment_length("bottom_wall", 0.0, 0.1)
    obs.add_velocity_profile((2.25, 1.5), 0, 0.5, 0.1)
    bump = cases.Problem(mesh=mesh, bcs=FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 10.0),
                         settings=s, obs=obs, params=InferenceParameterSet.a1_betaStar(),
                         nu=c["nu"], u_init=Vec3(0.3, 0.0, 0.0), p_init=0.0, k_init=1e-4,
                         omega_init=10.0)
    step = bfs_problem()
    obs = ObservationOperator()
    obs.add_reattachment_length("bottom_wall_down", 6.28, 0.5)
    obs.add_separation_point("bottom_wall_down", 0.0, 0.5)
    obs.add_separation_point("bottom_wall_up", 0.0, 0.5)
    obs.add_velocity_profile((4.0, 0.05), 0, 0.0, 0.01)
    step.obs = obs
    return {"cosine_bump": bump, "bfs": step}


def re180_problem():
    """The reference's committed periodic-channel regression
    (periodic_channel_re180.json, copied as data): makePeriodicChannel2D from
    the golden's mesh record, channelDefaults(1, 1, 10), U+ probes at x 0.5,
    the golden's settings with body force (1, 0, 0) and AMG pressure, init
    U 15, k 1, omega 10."""
    from cfdr_amd.config import (FlowBoundaryConditions, InferenceParameterSet,
                                 ObservationOperator, SolverSettings, Vec3)
    from cfdr_amd.mesh import Mesh
    g = load("ref_periodic_channel_re180.json")
    mp = g["mesh"]
    mesh = Mesh.make_periodic_channel_2d(mp["nx"], mp["ny"], mp["Lx"], mp["h"], mp["Re_tau"],
                                         mp["yplus_target"])
    mesh.compute_wall_distance()
    obs = ObservationOperator()
    for y in g["qoi"]["probe_y_over_delta"]:
        obs.add_velocity_profile((0.5, y, 0.0), 0, 0.0, 0.1)
    st = g["settings"]
    s = SolverSettings(max_iterations=st["max_iterations"], convergence_tol=st["convergence_tol"],
                       alpha_u=st["alpha_u"], alpha_p=st["alpha_p"], alpha_k=st["alpha_k"],
                       alpha_omega=st["alpha_omega"], inner_iterations=st["inner_iterations"],
                       inner_tolerance=st["inner_tolerance"], turb_start_iter=st["turb_start_iter"],
                       turb_update_interval=st["turb_update_interval"],
                       divergence_limit=st["divergence_limit"], body_force=Vec3(1.0, 0, 0))
    return cases.Problem(mesh=mesh, bcs=FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1.0, 10.0),
                         settings=s, obs=obs, params=InferenceParameterSet.a1_betaStar(),
                         nu=g["nu"], u_init=Vec3(15.0, 0.0, 0.0), p_init=0.0, k_init=1.0,
                         omega_init=10.0)


def bfs_lemoin_problem():
    """The reference's committed Le-Moin-Kim regression
    (tests/tools/record_goldens.cpp:463-497, golden data in
    ref_bfs_lemoin_uniform_inlet.json): makeBackwardFacingStep2D(40, 48, 24,
    18, 10, 22, 1, 6, 5100, 1) with a symmetry top (MSH2 fixture written by the
    reference's generator), bfsDefaults, Re_h 5100, Tu 5 %, AMG pressure,
    a reattachment length, 7 pressure taps, 7 skin-friction stations and a
    velocity profile at x 4."""
    from cfdr_amd.config import (FlowBoundaryConditions, InferenceParameterSet,
                                 ObservationOperator, SolverSettings, Vec3)
    from cfdr_amd.mesh import Mesh
    g = load("ref_bfs_lemoin_uniform_inlet.json")
    mesh = Mesh.load_binary(os.path.join(GOLDEN, "ref_bfs_lemoin_40x48.msh"))
    mesh.compute_wall_distance()
    nu, ub = g["nu"], 1.0
    k_in = 1.5 * (ub * 0.05) * (ub * 0.05)
    om_in = k_in / (nu * 100.0)
    st = g["settings"]
    s = SolverSettings(max_iterations=st["max_iterations"], convergence_tol=st["convergence_tol"],
                       divergence_limit=st["divergence_limit"], alpha_u=st["alpha_u"],
                       alpha_p=st["alpha_p"], alpha_k=st["alpha_k"], alpha_omega=st["alpha_omega"],
                       inner_iterations=st["inner_iterations"],
                       inner_tolerance=st["inner_tolerance"], turb_start_iter=st["turb_start_iter"],
                       turb_update_interval=st["turb_update_interval"])
    q = g["qoi"]
    obs = ObservationOperator()
    obs.add_reattachment_length("bottom_wall_down", 6.28, 0.5)
    for x in q["tap_x"]:
        obs.add_pressure_tap((x, 0.0), 0.0, 0.01, ub)
    for x in q["tap_x"]:
        obs.add_skin_friction("bottom_wall_down", (x, 0.0), 0.0, 1e-3, ub)
    for y in q["profile_y_at_x4"]:
        obs.add_velocity_profile((4.0, y), 0, 0.0, 0.01)
    return cases.Problem(mesh=mesh, bcs=FlowBoundaryConditions.channel_defaults(mesh, ub, k_in, om_in),
                         settings=s, obs=obs, params=InferenceParameterSet.a1_betaStar(), nu=nu,
                         u_init=Vec3(ub, 0.0, 0.0), p_init=0.0, k_init=k_in, omega_init=om_in)


def drag_problems():
    """The two problems of ref_drag.json: the flat plate observed by the drag
    of its plate and a velocity probe, and the backward-facing step observed
    by the drag of its step face and of the lower wall behind it and a
    velocity probe."""
    from cfdr_amd.config import ObservationOperator
    plate = plate_problem()
    obs = ObservationOperator()
    obs.add_drag("bottom", 0.004, 0.001, 1.0, 1.0)
    obs.add_velocity_profile((0.5, 0.01), 0, 0.9, 0.01)
    plate.obs = obs
    step = bfs_problem()
    obs = ObservationOperator()
    obs.add_drag("step_face", 0.0, 0.1, 1.0, 1.0)
    obs.add_drag("bottom_wall_down", 0.0, 0.01, 22.0, 1.0)
    obs.add_velocity_profile((4.0, 0.05), 0, 0.0, 0.01)
    step.obs = obs
    return {"plate": plate, "bfs": step}


def symmetry_channel():
    """1x32 geometric channel (ratio 1.15) whose top patch is a symmetry
    plane, 5000 capped iterations at tol 1e-6."""
    from cfdr_amd.config import FlowBoundaryConditions
    prob = cases.channel_problem(ny=32, ratio=1.15, n_probes=12, tol=1e-6, max_iterations=5000)
    mesh = prob.mesh
    for p in mesh.patches:
        if p.name == "top":
            p.type = "symmetry"
    mesh.compute_wall_distance()
    prob.bcs = FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1.0, 10.0)
    return prob


def max_rel(a, b):
    """Largest relative difference, measured against b (0 when both are empty)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def max_scaled(a, b):
    """Largest difference scaled by the field's own magnitude (for fields that
    cross zero, where a pointwise relative error is meaningless)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def sens_channel():
    """The periodic problem of the reference's warm-FD refinement test
    (rans-uq/tests/test_warm_fd_refinement.cpp:29-47): makePeriodicChannel2D(3, 4,
    1, 0.5, 20), channelDefaults(1, 1e-4, 1), nu 0.1, body force 0.2, upwind
    convection, tol 1e-9, no nuT floor window, one velocity probe."""
    from cfdr_amd.config import (FlowBoundaryConditions, InferenceParameterSet,
                                 ObservationOperator, SolverSettings, Vec3)
    from cfdr_amd.mesh import Mesh
    mesh = Mesh.make_periodic_channel_2d(3, 4, 1.0, 0.5, 20.0)
    mesh.compute_wall_distance()
    bcs = FlowBoundaryConditions.channel_defaults(mesh, 1.0, 1e-4, 1.0)
    s = SolverSettings(max_iterations=15000, convergence_tol=1e-9, nut_floor_iters=0,
                       momentum_convection="upwind", turbulence_convection="upwind",
                       body_force=Vec3(0.2, 0.0, 0.0))
    obs = ObservationOperator()
    obs.add_velocity_profile((0.43, 0.38, 0.0), 0, 0.0, 1.0)
    return cases.Problem(mesh=mesh, bcs=bcs, settings=s, obs=obs,
                         params=InferenceParameterSet.a1_betaStar(), nu=0.1,
                         u_init=Vec3(1.0, 0.0, 0.0), p_init=0.0, k_init=1e-4, omega_init=1.0)
