"""GPU parity of the gfx950 engine, called through the C-ABI
(include/rans_amd.h via cfdr_amd.engine), against the fixtures written by the
reference solver and against the CPU oracle on the same seeded inputs.

Tolerances: the converged/failed status must match exactly, and so must the
iteration count of capped and failed solves (converged solves: see
_iters_close); observation values (velocity probes) agree within
REL_OBS = 1e-6 relative, the bar the forward-solve contract sets. Field
comparisons of intermediate states use FIELD_TOL (scaled by the field's
magnitude). The engine follows the reference's FP64 operation order per cell
and per face, and only its dot products are summed in another (fixed) order,
so the observed differences are at round-off level (1e-15..1e-12); the stated
bounds are the contract, not the observation.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from cfdr_amd import cases, engine
from fixtures import (FIELD_NAMES, bfs_problem, couette_problem, crossings_problems,
                      bfs_lemoin_problem, drag_problems, inout_channel, load, max_rel,
                      max_scaled, plate_problem, re180_problem, small_channel, small_hill,
                      symmetry_channel)

pytestmark = pytest.mark.gpu

REL_OBS = 1e-6
FIELD_TOL = 1e-9


def _iters_close(got, want):
    """Iteration counts of converged solves. The engine's dot products are
    parallel reductions whose rounding differs from the reference's
    sequential sums, so the contraction measure can cross tol a monitor
    interval or two apart; capped and failed solves match exactly."""
    return abs(int(got) - int(want)) <= max(20, int(0.005 * want))


def _batch(prob, **kw):
    return engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                               prob.u_init, prob.p_init, prob.k_init, prob.omega_init, **kw)


def _oracle(prob):
    import oracle_py
    return oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs, prob.nu,
                                   prob.init_vector())


def _fields(res, i, nc):
    f = res.fields[i]
    out = {name: f[j * nc:(j + 1) * nc] for j, name in enumerate(FIELD_NAMES)}
    out["flux"] = f[10 * nc:]
    return out


def _check_row(res, i, row):
    assert res.statuses[i] == row["status"], (i, res.errors[i])
    if row["status"] == 0:
        assert _iters_close(res.iteration_counts[i], row["iters"])
    else:
        assert res.iteration_counts[i] == row["iters"]
    assert bool(res.stalled[i]) == row["stalled"]
    assert bool(res.convergence_mask[i]) == (row["status"] == 0)
    if row["status"] in (0, 1):
        assert max_rel(res.predictions[i], row["predictions"]) <= REL_OBS
        assert res.log_likelihoods[i] == pytest.approx(row["loglik"], rel=REL_OBS)
    else:
        assert np.all(np.isnan(res.predictions[i]))
        assert res.log_likelihoods[i] == row["loglik"]


def test_small_channel_rows_match_reference():
    d = load("ref_channel_small.json")
    prob = small_channel()
    res = _batch(prob).evaluate(np.asarray(d["thetas"]), fields=True)
    for i, row in enumerate(d["rows"]):
        _check_row(res, i, row)
        assert res.final_worst_measures[i] == pytest.approx(row["final_worst_measure"], rel=1e-4)
        assert res.contraction_rates[i] == pytest.approx(row["contraction_rate_u"], rel=1e-6)
    f = _fields(res, 0, prob.mesh.n_cells)
    for name in FIELD_NAMES + ["flux"]:
        assert max_scaled(f[name], d["rows"][0]["fields"][name]) <= FIELD_TOL, name


@pytest.mark.parametrize("K", ["1", "2", "6", "7", "30", "600"])
def test_intermediate_states_match_reference(K):
    """State after K iterations: initialisation, startup nuT floor, first
    turbulence updates, planar omega balance, Krylov warm starts."""
    d = load("ref_channel_steps.json")
    row = d["steps"][K]
    prob = small_channel(max_iterations=int(K))
    res = _batch(prob).evaluate(np.asarray([d["theta"]]), fields=True)
    assert res.statuses[0] == 1 and res.iteration_counts[0] == int(K)
    f = _fields(res, 0, prob.mesh.n_cells)
    for name in FIELD_NAMES + ["flux"]:
        assert max_scaled(f[name], row["fields"][name]) <= FIELD_TOL, name
    assert max_rel(res.predictions[0], row["predictions"]) <= REL_OBS


def test_status_edge_cases_match_reference():
    """ForwardModel.cpp:45-135 conventions on the GPU path."""
    e = load("ref_channel_edges.json")
    prob = small_channel()
    prob.settings.divergence_limit = 0.0
    res = _batch(prob).evaluate(np.asarray([[0.31, 0.09]]))
    _check_row(res, 0, e["divergence_limit_0"])
    assert res.final_worst_measures[0] == math.inf
    prob = small_channel(max_iterations=50)
    res = _batch(prob).evaluate(np.asarray([[0.31, 0.09]]))
    _check_row(res, 0, e["max_iterations_50"])
    prob = small_channel()
    prob.obs.observables[0].observed = 1e308
    res = _batch(prob).evaluate(np.asarray([[0.31, 0.09]]))
    row = e["likelihood_overflow"]
    assert (res.statuses[0], res.iteration_counts[0], res.log_likelihoods[0]) == \
        (row["status"], row["iters"], row["loglik"])
    assert np.all(np.isnan(res.predictions[0])) and not res.convergence_mask[0]
    assert "log-likelihood" in res.errors[0]
    prob = small_channel()
    res = _batch(prob).evaluate(np.asarray([[0.9, 0.09], [0.31, 0.09], [0.31, 5.0],
                                            [float("nan"), 0.09]]))
    for i in (0, 2, 3):
        _check_row(res, i, e["out_of_bounds"][0])
        assert "out of bounds" in res.errors[i]
    small = load("ref_channel_small.json")
    _check_row(res, 1, small["rows"][0])


def test_config1_channel_rows_match_reference():
    """1x128 geometric channel, tol 1e-8: default coefficients and the first
    three draws of the seeded 64-draw sweep."""
    d = load("ref_channel_c1.json")
    res = _batch(cases.channel_problem()).evaluate(np.asarray(d["thetas"]))
    for i, row in enumerate(d["rows"]):
        _check_row(res, i, row)


def test_three_column_channel_matches_oracle():
    """Three-column mesh: four faces per cell, distinct periodic neighbours."""
    prob = cases.channel_problem(ny=32, nx=3, ratio=1.15, n_probes=30, tol=1e-6,
                                 max_iterations=40000)
    theta = cases.channel_draws(3)
    res = _batch(prob).evaluate(theta)
    O = _oracle(prob)
    for i in range(len(theta)):
        o = O.evaluate(prob.coefficients(theta[i]))
        assert res.statuses[i] == o["status"]
        assert _iters_close(res.iteration_counts[i], o["iters"])
        assert max_rel(res.predictions[i], o["predictions"]) <= REL_OBS


@pytest.mark.parametrize("solver", ["PCG"])
def test_small_hill_matches_reference(solver):
    """30x16 periodic hill (curved, non-orthogonal faces, arena in HBM),
    400 capped iterations."""
    d = load(f"ref_hill_small_{solver.lower()}.json")
    prob = small_hill(solver)
    b = _batch(prob)
    res = b.evaluate(np.asarray([d["theta"]]), fields=True)
    _check_row(res, 0, d["row"])
    f = _fields(res, 0, prob.mesh.n_cells)
    for name in ("Ux", "Uy", "p", "k", "omega", "nuT", "flux"):
        assert max_scaled(f[name], d["row"]["fields"][name]) <= FIELD_TOL, name


@pytest.mark.parametrize("name", ["sst_no_limiter", "k_omega"])
def test_closure_variants_match_reference(name):
    v = load("ref_channel_variants.json")["variants"][name]
    prob = cases.channel_problem(ny=32, ratio=1.15, n_probes=10, tol=1e-6, max_iterations=40000)
    prob.obs.add_skin_friction("bottom", (0.5, 0.0), 0.005, 0.001, 15.0)
    prob.settings.sst_variant = v["sst_variant"]
    res = _batch(prob).evaluate(np.asarray([[0.31, 0.09]]))
    _check_row(res, 0, v["row"])


@pytest.mark.parametrize("scheme", ["limitedLinearUpwind", "linearUpwind", "upwind"])
def test_inlet_outlet_channel_matches_reference(scheme):
    """Inlet/outlet patches, pressure taps, skin friction, each scheme; the
    288-cell mesh keeps its arena in HBM."""
    d = load("ref_channel_inout.json")["schemes"][scheme]
    prob = inout_channel(d["momentum_convection"], d["turbulence_convection"])
    res = _batch(prob).evaluate(np.asarray(d["thetas"]), fields=True)
    for i, row in enumerate(d["rows"]):
        _check_row(res, i, row)
    f = _fields(res, 0, prob.mesh.n_cells)
    for name in FIELD_NAMES + ["flux"]:
        assert max_scaled(f[name], d["rows"][0]["fields"][name]) <= 1e-6, name


@pytest.mark.parametrize("mixed", [False, True])
def test_moving_wall_couette_matches_reference(mixed):
    """Moving top wall (WallMoving) on the graded 4x96 Couette channel (arena
    in HBM), PCG pressure: FP64 rows and fields follow the reference's; with
    FP32 Krylov plus FP64 defect correction the rows do. In Couette flow the
    wall-normal velocity and the pressure are zero up to round-off (|Uy| ~
    1e-8, |p| ~ 1e-7 against |Ux| ~ 2), so those two are scaled by the
    velocity scale max|Ux| (resp. its square) instead of their own size."""
    d = load("ref_couette.json")
    prob = couette_problem()
    res = _batch(prob, mixed_precision=mixed).evaluate(np.asarray(d["thetas"]), fields=True)
    for i, row in enumerate(d["rows"]):
        _check_row(res, i, row)
    if not mixed:
        f = _fields(res, 0, prob.mesh.n_cells)
        ref = d["rows"][0]["fields"]
        u = float(np.max(np.abs(ref["Ux"])))
        floor = {"Uy": u, "p": u * u}
        for name in FIELD_NAMES + ["flux"]:
            b = np.asarray(ref[name])
            scale = max(float(np.max(np.abs(b))), floor.get(name, 0.0))
            assert float(np.max(np.abs(f[name] - b))) <= 1e-6 * scale, name


@pytest.mark.parametrize("mixed", [False, True])
def test_free_stream_flat_plate_matches_reference(mixed):
    """Dirichlet velocity, k and omega on the free-stream top of the graded
    plate mesh (arena in HBM), PCG pressure: FP64 rows and fields follow the
    reference's; with FP32 Krylov plus FP64 defect correction the rows do."""
    d = load("ref_plate.json")
    prob = plate_problem()
    res = _batch(prob, mixed_precision=mixed).evaluate(np.asarray(d["thetas"]), fields=True)
    for i, row in enumerate(d["rows"]):
        _check_row(res, i, row)
    if not mixed:
        f = _fields(res, 0, prob.mesh.n_cells)
        for name in FIELD_NAMES + ["flux"]:
            assert max_scaled(f[name], d["rows"][0]["fields"][name]) <= 1e-6, name


@pytest.mark.parametrize("grid_blocks", [1, 4])
def test_backward_facing_step_matches_reference(grid_blocks):
    """The reference's backward-facing-step mesh (744 cells: step face,
    re-entrant corner, symmetry top, inlet and outlet; arena in HBM), 600
    capped iterations, one and four workgroups per draw: the row and every
    field follow the reference's. With FP32 Krylov plus FP64 defect
    correction the capped row keeps its status and count, and its probes stay
    within REL_OBS of the FP64 reference scaled by the probes' magnitude
    (a capped state whose skin-friction probes cross zero; measured 1.5e-7)."""
    d = load("ref_bfs.json")
    prob = bfs_problem()
    b = _batch(prob, grid_blocks=grid_blocks)
    res = b.evaluate(np.asarray([d["theta"]]), fields=True)
    assert b.last_timing()["grid_blocks"] == grid_blocks
    _check_row(res, 0, d["row"])
    f = _fields(res, 0, prob.mesh.n_cells)
    for name in FIELD_NAMES + ["flux"]:
        assert max_scaled(f[name], d["row"]["fields"][name]) <= 1e-6, name
    if grid_blocks == 1:
        res = _batch(prob, mixed_precision=True).evaluate(np.asarray([d["theta"]]))
        assert res.statuses[0] == d["row"]["status"]
        assert res.iteration_counts[0] == d["row"]["iters"]
        assert max_scaled(res.predictions[0], d["row"]["predictions"]) <= REL_OBS


@pytest.mark.parametrize("case", ["cosine_bump", "bfs"])
def test_separation_and_reattachment_match_reference(case):
    """Separation-point and reattachment-length observations on the cosine
    bump (arena in HBM) and the backward-facing step, capped runs: the rows
    follow the reference's, these observations included."""
    d = load("ref_wall_crossings.json")[case]
    prob = crossings_problems()[case]
    res = _batch(prob).evaluate(np.asarray(d["thetas"]))
    for i, row in enumerate(d["rows"]):
        _check_row(res, i, row)
    if case == "bfs":
        # four workgroups per draw in FP64 evaluate the crossings in the
        # multi-workgroup kernel; in mixed precision such a problem runs on
        # one workgroup per draw, with the capped status and count kept and
        # probes within REL_OBS of the FP64 reference, scaled by their
        # magnitude (measured 2.3e-7)
        b = _batch(prob, grid_blocks=4)
        res = b.evaluate(np.asarray(d["thetas"]))
        assert b.last_timing()["grid_blocks"] == 4
        _check_row(res, 0, d["rows"][0])
        b = _batch(prob, grid_blocks=4, mixed_precision=True)
        res = b.evaluate(np.asarray(d["thetas"]))
        assert b.last_timing()["grid_blocks"] == 1
        assert res.statuses[0] == d["rows"][0]["status"]
        assert res.iteration_counts[0] == d["rows"][0]["iters"]
        assert max_scaled(res.predictions[0], d["rows"][0]["predictions"]) <= REL_OBS


@pytest.mark.parametrize("case", ["plate", "bfs"])
def test_drag_matches_reference(case):
    """Drag observations on the flat plate (converged) and on the step face
    and lower wall of the backward-facing step (capped): the rows follow the
    reference's, the drag included. A problem with drag observations runs
    one workgroup per draw even when four are asked for."""
    d = load("ref_drag.json")[case]
    prob = drag_problems()[case]
    res = _batch(prob).evaluate(np.asarray(d["thetas"]))
    for i, row in enumerate(d["rows"]):
        _check_row(res, i, row)
    if case == "bfs":
        b = _batch(prob, grid_blocks=4)
        res = b.evaluate(np.asarray(d["thetas"]))
        assert b.last_timing()["grid_blocks"] == 1
        _check_row(res, 0, d["rows"][0])


def test_reference_periodic_channel_golden_amg():
    """The reference's committed periodic-channel regression (4x56, AMG
    pressure, 4,080 iterations, recorded with another compiler and libm): the
    device AMG gives its status and stalled flag, the parity bar's iteration
    count and U+ probes within REL_OBS."""
    g = load("ref_periodic_channel_re180.json")
    prob = re180_problem()
    res = _batch(prob).evaluate(np.asarray([[0.31, 0.09]]))
    r = g["result"]
    print("re180 golden", res.statuses[0], res.iteration_counts[0], r["final_iter"], flush=True)
    assert res.statuses[0] == 0 and bool(res.stalled[0]) == r["stalled"]
    assert _iters_close(res.iteration_counts[0], r["final_iter"])
    assert max_rel(res.predictions[0], g["qoi"]["u_plus_probes"]) <= REL_OBS


def test_reference_bfs_golden_amg():
    """The reference's committed Le-Moin-Kim backward-facing-step regression
    (2,976 cells, AMG pressure, converged in 4,060 iterations, recorded with
    another compiler and libm): the device AMG (Jacobi-smoothed) converges
    too, at the parity bar for converged solves, and each group of
    observations (reattachment length, pressure taps, skin friction, the
    velocity profile) stays within REL_OBS of the group's size (observed:
    the same 4,060 iterations and 1e-12)."""
    g = load("ref_bfs_lemoin_uniform_inlet.json")
    prob = bfs_lemoin_problem()
    res = _batch(prob).evaluate(np.asarray([[0.31, 0.09]]))
    its = g["result"]["final_iter"]
    print("bfs golden", res.statuses[0], res.iteration_counts[0], its, flush=True)
    assert res.statuses[0] == 0 and not res.stalled[0]
    assert _iters_close(res.iteration_counts[0], its)
    q, p = g["qoi"], np.asarray(res.predictions[0])
    n = len(q["tap_x"])
    groups = {"x_reattachment": (p[:1], [q["x_reattachment"]]),
              "cp": (p[1:1 + n], q["cp_taps"]), "cf": (p[1 + n:1 + 2 * n], q["cf_taps"]),
              "u": (p[1 + 2 * n:], q["u_profile_x4"])}
    for name, (got, want) in groups.items():
        err = max_scaled(got, want)
        print(name, err, flush=True)
        assert err <= REL_OBS, name


def test_reference_couette_golden_amg():
    """The reference's committed Couette regression (its AMG pressure
    solver, recorded with another compiler and libm): the device AMG
    (Jacobi-smoothed) reaches the same status, the parity bar's iteration
    count and probes within REL_OBS."""
    g = load("ref_couette_re171_ny96.json")
    prob = couette_problem(pressure_solver="AMG")
    res = _batch(prob).evaluate(np.asarray([[0.31, 0.09]]))
    its = g["result"]["final_iter"]
    q = g["qoi"]
    err = max_rel(res.predictions[0], [q["cf_probe"]] + q["u_probes"])
    print("couette golden", res.statuses[0], res.iteration_counts[0], its, err, flush=True)
    assert res.statuses[0] == 0 and not res.stalled[0]
    assert _iters_close(res.iteration_counts[0], its)
    assert err <= REL_OBS


@pytest.mark.parametrize("mixed", [False, True])
def test_wall_normal_body_force_matches_reference(mixed):
    """A y body force: the 1x32 periodic channel (arena in LDS) under
    (1, 0.25) and the 288-cell developing channel (arena in HBM) under
    (0, -0.2). FP64 rows and fields follow the reference; with FP32 Krylov
    plus FP64 defect correction the rows do (status, iterations, probes and
    log-likelihood at REL_OBS)."""
    d = load("ref_channel_body_force.json")
    c = d["periodic"]
    prob = small_channel()
    prob.settings.body_force.y = c["body_force"][1]
    res = _batch(prob, mixed_precision=mixed).evaluate(np.asarray([c["theta"]]), fields=True)
    _check_row(res, 0, c["row"])
    if not mixed:
        f = _fields(res, 0, prob.mesh.n_cells)
        for name in ("Ux", "Uy", "p", "k", "omega", "nuT", "flux"):
            assert max_scaled(f[name], c["row"]["fields"][name]) <= FIELD_TOL, name
    io = d["inout"]
    prob = inout_channel("upwind", "upwind")
    prob.settings.body_force.y = io["body_force"][1]
    res = _batch(prob, mixed_precision=mixed).evaluate(np.asarray(io["thetas"]), fields=True)
    for i, row in enumerate(io["rows"]):
        _check_row(res, i, row)
    if not mixed:
        f = _fields(res, 0, prob.mesh.n_cells)
        for name in FIELD_NAMES + ["flux"]:
            assert max_scaled(f[name], io["rows"][0]["fields"][name]) <= 1e-6, name


def test_symmetry_patch_matches_reference():
    d = load("ref_channel_symmetry.json")
    prob = symmetry_channel()
    res = _batch(prob).evaluate(np.asarray([d["theta"]]), fields=True)
    _check_row(res, 0, d["row"])
    f = _fields(res, 0, prob.mesh.n_cells)
    for name in FIELD_NAMES + ["flux"]:
        assert max_scaled(f[name], d["row"]["fields"][name]) <= FIELD_TOL, name


def test_mixed_precision_matches_fp64_reference():
    """FP32 Krylov iterations with FP64 defect correction: same converged
    flags as the FP64 reference, probes within REL_OBS of it."""
    for prob, fixture in ((small_channel(), "ref_channel_small.json"),
                          (cases.channel_problem(), "ref_channel_c1.json")):
        d = load(fixture)
        res = _batch(prob, mixed_precision=True).evaluate(np.asarray(d["thetas"]))
        for i, row in enumerate(d["rows"]):
            assert res.statuses[i] == row["status"], (fixture, i)
            assert _iters_close(res.iteration_counts[i], row["iters"])
            assert max_rel(res.predictions[i], row["predictions"]) <= REL_OBS


def _rows_of(d):
    return (np.asarray(d["thetas"]), [r["status"] for r in d["rows"]],
            [r["iters"] for r in d["rows"]], [r["predictions"] for r in d["rows"]])


def test_mixed_precision_block_paths_match_fp64_reference():
    """FP32 block Krylov with FP64 defect correction, beyond the
    wave-resident path: 512-thread workgroups on the 128-cell channel (arena
    in LDS) and the 288-cell inlet/outlet channel (arena in HBM, one and four
    workgroups per draw), against the FP64 reference at REL_OBS."""
    th, st, its, pred = _rows_of(load("ref_channel_c1.json"))
    res = _batch(cases.channel_problem(), mixed_precision=True, block_size=512).evaluate(th)
    for i in range(len(th)):
        assert res.statuses[i] == st[i] and _iters_close(res.iteration_counts[i], its[i])
        assert max_rel(res.predictions[i], pred[i]) <= REL_OBS
    d = load("ref_channel_inout.json")["schemes"]["limitedLinearUpwind"]
    prob = inout_channel(d["momentum_convection"], d["turbulence_convection"])
    th, st, its, pred = _rows_of(d)
    for g in (1, 4):
        b = _batch(prob, mixed_precision=True, grid_blocks=g)
        res = b.evaluate(th)
        assert b.last_timing()["grid_blocks"] == g
        for i in range(len(th)):
            assert res.statuses[i] == st[i] and _iters_close(res.iteration_counts[i], its[i]), g
            assert max_rel(res.predictions[i], pred[i]) <= REL_OBS, g


def test_mixed_precision_amg_hill_matches_fp64_reference():
    """FP32 AMG-preconditioned CG (level-0 smoothing and Krylov vectors in
    FP32, coarser levels FP64) on the converged 30x16 hills, one and eight
    workgroups per draw: the reference's statuses and iteration counts, and
    probes within REL_OBS of the reference's AMG (measured 4.7e-9)."""
    d = load("ref_hill_small_converged.json")
    th, st, its, pred = _rows_of(d)
    prob = cases.hill_problem(nx=30, ny=16)
    for g in (1, 8):
        res = _batch(prob, mixed_precision=True, grid_blocks=g).evaluate(th)
        for i in range(len(th)):
            assert res.statuses[i] == st[i] == 0
            assert _iters_close(res.iteration_counts[i], its[i])
            assert max_rel(res.predictions[i], pred[i]) <= REL_OBS


def test_amg_matches_pcg_at_tight_tolerance():
    """The device AMG-preconditioned CG solves the same pressure systems as
    Jacobi-PCG: with tight inner tolerances the two give the same state."""
    out = {}
    for solver in ("PCG", "AMG"):
        prob = cases.hill_problem(nx=30, ny=16, max_iterations=40, pressure_solver=solver)
        prob.settings.inner_tolerance = 1e-11
        prob.settings.inner_iterations = 3000
        res = _batch(prob).evaluate(np.asarray([[0.31, 0.09, 0.85, 0.5]]), fields=True)
        out[solver] = (res, _fields(res, 0, prob.mesh.n_cells))
    for name in ("Ux", "Uy", "p", "k", "omega", "flux"):
        assert max_scaled(out["AMG"][1][name], out["PCG"][1][name]) <= 1e-9, name
    assert out["AMG"][0].krylov_iterations[0][2] * 5 < out["PCG"][0].krylov_iterations[0][2]


def test_converged_hill_amg_matches_reference():
    """Converged 30x16 hills: the reference's AMG (Gauss-Seidel smoothing)
    against the device AMG (damped-Jacobi smoothing, coarser aggregates):
    both reach the same fixed point, with the reference's iteration counts and
    probes within REL_OBS (measured 4.6e-9)."""
    d = load("ref_hill_small_converged.json")
    prob = cases.hill_problem(nx=30, ny=16)
    res = _batch(prob).evaluate(np.asarray(d["thetas"]))
    for i, row in enumerate(d["rows"]):
        assert res.statuses[i] == row["status"] == 0
        assert _iters_close(res.iteration_counts[i], row["iters"])
        assert max_rel(res.predictions[i], row["predictions"]) <= REL_OBS


@pytest.mark.parametrize("solver", ["PCG", "AMG"])
def test_multi_workgroup_draws_match_single_workgroup(solver):
    """Four workgroups per draw (grid barriers, rank-ordered reductions)
    against one workgroup per draw on the 30x16 hill, 400 capped iterations."""
    prob = small_hill(solver)
    theta = np.asarray([[0.31, 0.09, 0.85, 0.5], [0.3, 0.095, 0.9, 0.55]])
    one = _batch(prob, grid_blocks=1).evaluate(theta, fields=True)
    b4 = _batch(prob, grid_blocks=4)
    four = b4.evaluate(theta, fields=True)
    assert b4.last_timing()["grid_blocks"] == 4
    nc = prob.mesh.n_cells
    for i in range(len(theta)):
        assert four.statuses[i] == one.statuses[i] and four.iteration_counts[i] == one.iteration_counts[i]
        assert max_rel(four.predictions[i], one.predictions[i]) <= REL_OBS
        f4, f1 = _fields(four, i, nc), _fields(one, i, nc)
        for name in ("Ux", "Uy", "p", "k", "omega"):
            assert max_scaled(f4[name], f1[name]) <= 1e-6, name


def test_multi_workgroup_converged_hill_matches_reference():
    d = load("ref_hill_small_converged.json")
    prob = cases.hill_problem(nx=30, ny=16)
    res = _batch(prob, grid_blocks=8).evaluate(np.asarray(d["thetas"]))
    for i, row in enumerate(d["rows"]):
        assert res.statuses[i] == row["status"] == 0
        assert _iters_close(res.iteration_counts[i], row["iters"])
        assert max_rel(res.predictions[i], row["predictions"]) <= REL_OBS


def test_batch_results_independent_of_batch_composition():
    """Each row is bitwise the same alone, inside a batch and in another order
    (test_batch_determinism.cpp:49-63)."""
    prob = small_channel()
    theta = cases.channel_draws(6)
    b = _batch(prob)
    full = b.evaluate(theta)
    rev = b.evaluate(theta[::-1].copy())
    for i in range(len(theta)):
        alone = b.evaluate(theta[i:i + 1])
        for other, j in ((alone, 0), (rev, len(theta) - 1 - i)):
            assert other.statuses[j] == full.statuses[i]
            assert other.iteration_counts[j] == full.iteration_counts[i]
            assert np.array_equal(other.predictions[j], full.predictions[i])
            assert other.log_likelihoods[j] == full.log_likelihoods[i]


def test_block_sizes_agree():
    """Launch shape changes only the reduction grouping."""
    prob = small_channel()
    theta = cases.channel_draws(2)
    base = _batch(prob, block_size=64).evaluate(theta)
    for bs in (128, 256):
        r = _batch(prob, block_size=bs).evaluate(theta)
        assert np.array_equal(r.status_codes, base.status_codes)
        assert max_rel(r.predictions, base.predictions) <= REL_OBS


def test_empty_batch_and_argument_errors():
    prob = small_channel()
    b = _batch(prob)
    r = b.evaluate(np.zeros((0, 2)))
    assert r.predictions.shape == (0, prob.obs.n_obs())
    with pytest.raises(ValueError, match="workers"):
        b.evaluate(np.asarray([[0.31, 0.09]]), workers=0)
    with pytest.raises(ValueError, match="n_parameters"):
        b.evaluate(np.asarray([[0.31, 0.09, 0.5]]))


def test_sweep_draws_match_oracle():
    """Eight draws of the seeded channel sweep on the small channel against the
    oracle run on the same inputs."""
    prob = small_channel()
    theta = cases.channel_draws(8)
    res = _batch(prob).evaluate(theta)
    O = _oracle(prob)
    for i in range(len(theta)):
        o = O.evaluate(prob.coefficients(theta[i]))
        assert res.statuses[i] == o["status"]
        assert _iters_close(res.iteration_counts[i], o["iters"])
        assert max_rel(res.predictions[i], o["predictions"]) <= REL_OBS
        assert res.krylov_iterations[i].tolist() == o["krylov_iters"]
