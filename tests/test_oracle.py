"""Pins the CPU oracle (oracle/rans_oracle.c) before it is trusted as the
checker of the HIP engine.

* known answers and closed forms for the pointwise SST-2003 functions
  (core/tests/test_sst_production_2003.cpp:54-141 and the published
  Menter-Kuntz-Langtry 2003 definitions);
* exact solutions of small linear systems for every Krylov/AMG solver
  (core/tests/test_linear_scratch.cpp:113-119,202-229) and a random SPD
  Laplacian against numpy;
* the reference's own committed regression goldens for the periodic channel
  and the moving-wall Couette flow
  (tests/golden/constrained_omega/periodic_channel_re180.json and
  couette_re171_ny96.json of the reference, copied as data);
* fixtures written by the reference solver compiled from its sources
  (tests/golden/make_fixtures.py): full solves, intermediate states after K
  iterations, status edge cases, the config-1 channel and a small hill;
* physics invariants of the converged channel (wall shear balances the body
  force, streamwise invariance, one column equals three columns).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle_py
from cfdr_amd import cases
from cfdr_amd.config import COEFF_DEFAULTS
from cfdr_amd.mesh import Mesh
from fixtures import (FIELD_NAMES, bfs_problem, couette_problem, crossings_problems,
                      drag_problems, inout_channel, load, max_rel, plate_problem, small_channel,
                      small_hill, symmetry_channel)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _c(coeffs=COEFF_DEFAULTS):
    a = np.ascontiguousarray(coeffs, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


# --------------------------------------------------------------- SST pointwise
def test_production_omega_known_answers():
    """core/tests/test_sst_production_2003.cpp:62-103."""
    L = oracle_py.lib()
    keep, c = _c()
    b_star = COEFF_DEFAULTS[8]
    # limiter inactive: exactly S^2
    assert L.or_sst_pomega(c, 0.01, 1.0, 1.0, 100.0) == pytest.approx(1.0, rel=1e-14)
    # limiter active: 10 b* k w / nuT = 0.9
    p = L.or_sst_pomega(c, 1.0, 100.0, 1.0, 1.0)
    assert p == pytest.approx(10 * b_star, rel=1e-14)
    assert p < 1e-3 * 100.0 ** 2
    # bounded by S^2 across a sweep
    for nut in (1e-12, 1e-6, 1e-2, 0.5, 3.0):
        for s in (0.0, 0.3, 3.0, 300.0):
            for k in (1e-8, 1e-2, 1.0, 50.0):
                for w in (1e-3, 1.0, 1e4):
                    assert L.or_sst_pomega(c, nut, s, k, w) <= s * s + 1e-14
    # nuT -> 0 selects S^2
    assert L.or_sst_pomega(c, 0.0, 2.0, 0.5, 200.0) == pytest.approx(4.0, rel=1e-14)


def test_production_omega_matches_k_production():
    """productionOmega == production()/nuT for nuT > 0
    (test_sst_production_2003.cpp:105-119), and the limited k production is
    min(nuT S^2, 10 b* k w)."""
    L = oracle_py.lib()
    keep, c = _c()
    for nut, s, k, w in ((0.01, 1.0, 1.0, 100.0), (1.0, 100.0, 1.0, 1.0), (0.2, 5.0, 2.0, 30.0)):
        by_min = L.or_sst_pomega(c, nut, s, k, w)
        by_ratio = L.or_sst_pk(c, nut, s, k, w) / nut
        assert by_min == pytest.approx(by_ratio, rel=1e-12)
    assert L.or_sst_pk(c, 0.01, 1.0, 1.0, 100.0) == pytest.approx(0.01, rel=1e-15)
    assert L.or_sst_pk(c, 1.0, 100.0, 1.0, 1.0) == pytest.approx(0.9, rel=1e-15)


def test_blending_functions_against_published_definitions():
    """F1 = tanh(arg1^4), F2 = tanh(arg2^2), nuT = a1 k / max(a1 w, S F2)
    (Menter, Kuntz & Langtry 2003) in both branches of every max/min."""
    L = oracle_py.lib()
    keep, c = _c()
    bs, sw2, a1 = COEFF_DEFAULTS[8], COEFF_DEFAULTS[5], COEFF_DEFAULTS[9]
    nu = 1.0 / 182.088
    for k, w, y, cd in ((1.0, 10.0, 0.01, 1e-3), (1e-3, 1e4, 1e-4, 0.0), (2.0, 0.5, 0.8, 5.0),
                        (0.3, 80.0, 0.05, -2.0)):
        t1 = math.sqrt(k) / (bs * w * y)
        t2 = 500 * nu / (y * y * w)
        t3 = 4 * sw2 * k / (max(cd, 1e-10) * y * y)
        f1 = math.tanh(min(max(t1, t2), t3) ** 4)
        f2 = math.tanh(max(2 * t1, t2) ** 2)
        assert L.or_sst_f1(c, k, w, y, nu, cd) == pytest.approx(f1, rel=1e-14, abs=1e-300)
        assert L.or_sst_f2(c, k, w, y, nu) == pytest.approx(f2, rel=1e-14, abs=1e-300)
    # near the wall both blends are 1, far from it with tiny k both are ~0
    assert L.or_sst_f1(c, 1.0, 1e3, 1e-5, nu, 1e-3) == 1.0
    assert L.or_sst_f2(c, 1e-8, 1e3, 5.0, nu) < 1e-6
    # nuT: S = 0 gives k/w exactly; a shear-limited state gives a1 k / (S F2)
    assert L.or_sst_nut(c, 0, 2.0, 4.0, 0.0, 1.0) == pytest.approx(0.5, rel=1e-15)
    assert L.or_sst_nut(c, 0, 1.0, 1.0, 10.0, 1.0) == pytest.approx(a1 / 10.0, rel=1e-15)
    # NoLimiter variant is plain k / w
    assert L.or_sst_nut(c, 1, 3.0, 2.0, 1e3, 1.0) == 1.5


# ------------------------------------------------------------- linear solvers
def _ldu_solve(solver, n, own, nbr, diag, upper, lower, b, tol=1e-12, max_iter=500, x0=None):
    L = oracle_py.lib()
    own = np.ascontiguousarray(own, np.int32)
    nbr = np.ascontiguousarray(nbr, np.int32)
    arrs = [np.ascontiguousarray(a, np.float64) for a in (diag, upper, lower, b)]
    x = np.zeros(n) if x0 is None else np.array(x0, float)
    it = C.c_int(0)
    r0 = C.c_double(0)
    r1 = C.c_double(0)
    rc = L.or_linear_solve(solver, n, len(own), own.ctypes.data_as(_ip), nbr.ctypes.data_as(_ip),
                           *[a.ctypes.data_as(_dp) for a in arrs], x.ctypes.data_as(_dp),
                           max_iter, tol, C.byref(it), C.byref(r0), C.byref(r1))
    return rc, x, it.value


@pytest.mark.parametrize("solver", [0, 1, 2])
def test_diagonal_system_exact(solver):
    """diag(2,4) x = (6,-8) -> x = (3,-2) (core/tests/test_linear_scratch.cpp:113-119)."""
    rc, x, _ = _ldu_solve(solver, 2, [], [], [2.0, 4.0], [], [], [6.0, -8.0])
    assert rc == 1
    np.testing.assert_allclose(x, [3.0, -2.0], rtol=1e-12)


@pytest.mark.parametrize("solver", [0, 1, 2])
def test_tridiagonal_system_exact(solver):
    """[[2,-1,0],[-1,2,-1],[0,-1,2]] x = (0,2,0)... exact (1,2,-1) is the
    reference's CSR case (test_linear_scratch.cpp:202-229): b = A (1,2,-1)."""
    x_true = np.array([1.0, 2.0, -1.0])
    A = np.array([[4.0, -1.0, 0.0], [-1.0, 4.0, -1.0], [0.0, -1.0, 4.0]])
    b = A @ x_true
    rc, x, _ = _ldu_solve(solver, 3, [0, 1], [1, 2], [4.0, 4.0, 4.0], [-1.0, -1.0], [-1.0, -1.0], b)
    assert rc == 1
    np.testing.assert_allclose(x, x_true, rtol=1e-10)


@pytest.mark.parametrize("solver", [0, 1, 2])
def test_random_spd_laplacian_against_numpy(solver):
    """Seeded 2D 12x9 five-point Laplacian (diagonally dominant): every solver
    reaches numpy's solution to its tolerance."""
    rng = np.random.default_rng(7)
    nx, ny = 12, 9
    n = nx * ny
    own, nbr = [], []
    for j in range(ny):
        for i in range(nx):
            c = j * nx + i
            if i + 1 < nx:
                own.append(c)
                nbr.append(c + 1)
            if j + 1 < ny:
                own.append(c)
                nbr.append(c + nx)
    off = -rng.uniform(0.5, 1.5, len(own))
    diag = np.full(n, 0.1)
    for f, (o, m) in enumerate(zip(own, nbr)):
        diag[o] -= off[f]
        diag[m] -= off[f]
    A = np.diag(diag)
    for f, (o, m) in enumerate(zip(own, nbr)):
        A[o, m] = A[m, o] = off[f]
    b = rng.standard_normal(n)
    rc, x, iters = _ldu_solve(solver, n, own, nbr, diag, off, off, b, tol=1e-12, max_iter=2000)
    assert rc == 1 and iters > 0
    x_np = np.linalg.solve(A, b)
    assert np.max(np.abs(x - x_np)) <= 1e-8 * np.max(np.abs(x_np))


# ------------------------------------------------- mesh generators (geometry)
def test_generator_meshes_bitwise_equal_reference():
    """makePeriodicChannel2D / makeCurvedChannelPeriodic2D geometry written by
    the reference (core/src/Mesh.cpp:511-741)."""
    from cfdr_amd.mesh import hill_bottom
    ref = load("ref_meshes.json")
    a = ref["periodic_channel_4x8"]["args"]
    m = Mesh.make_periodic_channel_2d(*a)
    built = {"periodic_channel_4x8": m}
    nx = 6
    xs = [9.0 * i / nx for i in range(nx + 1)]
    ys = [hill_bottom(x) for x in xs]
    ys[nx] = ys[0]
    built["hill_6x4"] = Mesh.make_curved_channel_periodic_2d(xs, ys, 3.035, 4, 5600.0, 1.0)
    for name, mesh in built.items():
        want = ref[name]["mesh"]
        for key, values in want.items():
            got = np.asarray(getattr(mesh, key)).ravel()
            assert np.array_equal(got, np.asarray(values, dtype=got.dtype)), (name, key)


# ------------------------------------------------------ reference's own golden
def test_reference_periodic_channel_golden():
    """The reference's committed regression (4x56 periodic channel, AMG p',
    tol 1e-4). Recorded with another compiler and libm, so iteration count and
    status must match exactly and values to the last few digits."""
    g = load("ref_periodic_channel_re180.json")
    from cfdr_amd.config import FlowBoundaryConditions, ObservationOperator, SolverSettings, Vec3
    mp = g["mesh"]
    m = Mesh.make_periodic_channel_2d(mp["nx"], mp["ny"], mp["Lx"], mp["h"], mp["Re_tau"],
                                      mp["yplus_target"])
    m.compute_wall_distance()
    bcs = FlowBoundaryConditions.channel_defaults(m, 1.0, 1.0, 10.0)
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
    s.validate()
    P = oracle_py.OracleProblem(m, bcs, s, obs, g["nu"], [15.0, 0.0, 0.0, 1.0, 10.0])
    r = P.evaluate(COEFF_DEFAULTS, fields=True)
    res = g["result"]
    assert r["status"] == 0 and r["iters"] == res["final_iter"]
    assert r["stalled"] == res["stalled"]
    assert max_rel(r["predictions"], g["qoi"]["u_plus_probes"]) < 1e-9
    assert r["worst"] == pytest.approx(res["final_worst_measure"], rel=1e-5)
    assert r["rate"] == pytest.approx(res["contraction_rate_u"], rel=1e-9)
    norms = res["field_norms"]
    assert float(np.sum(r["Ux"] ** 2)) == pytest.approx(norms["sumsq_ux"], rel=1e-9)
    assert float(np.sum(r["k"] ** 2)) == pytest.approx(norms["sumsq_k"], rel=1e-9)
    assert float(np.sum(r["omega"] ** 2)) == pytest.approx(norms["sumsq_omega"], rel=1e-9)
    assert float(np.sum(r["nuT"] ** 2)) == pytest.approx(norms["sumsq_nut"], rel=1e-9)


# ------------------------------------------- fixtures of the compiled reference
def _oracle(prob):
    return oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs, prob.nu,
                                   prob.init_vector())


def _check_row(r, row):
    assert r["status"] == row["status"]
    assert r["iters"] == row["iters"]
    assert r["stalled"] == row["stalled"]
    if row["status"] in (0, 1):
        assert max_rel(r["predictions"], row["predictions"]) == 0.0
        assert r["loglik"] == row["loglik"]
        assert r["worst"] == row["final_worst_measure"]
        rate = row["contraction_rate_u"]
        assert (rate is None and math.isnan(r["rate"])) or r["rate"] == rate
    else:
        assert np.all(np.isnan(r["predictions"]))
        assert r["loglik"] == row["loglik"]


def test_small_channel_full_solves():
    d = load("ref_channel_small.json")
    prob = small_channel()
    P = _oracle(prob)
    for i, (theta, row) in enumerate(zip(d["thetas"], d["rows"])):
        r = P.evaluate(prob.coefficients(theta), fields="fields" in row)
        _check_row(r, row)
        if "fields" in row:
            for name in FIELD_NAMES + ["flux"]:
                assert np.array_equal(r[name], np.asarray(row["fields"][name])), name


@pytest.mark.parametrize("K", ["1", "2", "6", "7", "30", "600"])
def test_small_channel_intermediate_states(K):
    """State after K outer iterations: covers initialisation, the startup
    nuT floor, the first turbulence updates and the planar omega balance."""
    d = load("ref_channel_steps.json")
    row = d["steps"][K]
    prob = small_channel(max_iterations=int(K))
    r = _oracle(prob).evaluate(prob.coefficients(d["theta"]), fields=True)
    assert r["status"] == row["status"] == 1 and r["iters"] == row["iters"] == int(K)
    for name in FIELD_NAMES + ["flux"]:
        assert np.array_equal(r[name], np.asarray(row["fields"][name])), name


def test_status_edge_cases():
    """ForwardModel.cpp:45-135 conventions: zero divergence limit, iteration
    cap, likelihood overflow, out-of-bounds coefficients."""
    e = load("ref_channel_edges.json")
    prob = small_channel()
    prob.settings.divergence_limit = 0.0
    _check_row(_oracle(prob).evaluate(prob.coefficients([0.31, 0.09])), e["divergence_limit_0"])
    prob = small_channel(max_iterations=50)
    _check_row(_oracle(prob).evaluate(prob.coefficients([0.31, 0.09])), e["max_iterations_50"])
    prob = small_channel()
    prob.obs.observables[0].observed = 1e308
    r = _oracle(prob).evaluate(prob.coefficients([0.31, 0.09]))
    row = e["likelihood_overflow"]
    assert (r["status"], r["iters"], r["loglik"]) == (row["status"], row["iters"], row["loglik"])


@pytest.mark.parametrize("solver", ["AMG", "PCG"])
def test_small_hill_capped(solver):
    """30x16 periodic hill (curved, non-orthogonal), 400 capped iterations."""
    d = load(f"ref_hill_small_{solver.lower()}.json")
    prob = small_hill(solver)
    r = _oracle(prob).evaluate(prob.coefficients(d["theta"]), fields=True)
    row = d["row"]
    _check_row(r, row)
    for name in FIELD_NAMES + ["flux"]:
        assert np.array_equal(r[name], np.asarray(row["fields"][name])), name


def test_config1_channel_rows():
    """1x128 geometric channel at tol 1e-8 (the single-solve workload), the
    default coefficients and the first draw of the 64-draw sweep."""
    d = load("ref_channel_c1.json")
    prob = cases.channel_problem()
    P = _oracle(prob)
    for theta, row in list(zip(d["thetas"], d["rows"]))[:2]:
        _check_row(P.evaluate(prob.coefficients(theta)), row)


@pytest.mark.parametrize("name", ["sst_no_limiter", "k_omega"])
def test_closure_variants(name):
    """SST NoLimiter and k-omega variants, with a skin-friction observation."""
    v = load("ref_channel_variants.json")["variants"][name]
    prob = cases.channel_problem(ny=32, ratio=1.15, n_probes=10, tol=1e-6, max_iterations=40000)
    prob.obs.add_skin_friction("bottom", (0.5, 0.0), 0.005, 0.001, 15.0)
    prob.settings.sst_variant = v["sst_variant"]
    _check_row(_oracle(prob).evaluate(prob.coefficients([0.31, 0.09])), v["row"])


@pytest.mark.parametrize("scheme", ["limitedLinearUpwind", "linearUpwind", "upwind"])
def test_inlet_outlet_channel(scheme):
    """Developing channel with inlet/outlet patches (outlet pressure, no p'
    pin), pressure taps and skin friction, for each convection scheme."""
    d = load("ref_channel_inout.json")["schemes"][scheme]
    prob = inout_channel(d["momentum_convection"], d["turbulence_convection"])
    P = _oracle(prob)
    for i, (theta, row) in enumerate(zip(d["thetas"], d["rows"])):
        r = P.evaluate(prob.coefficients(theta), fields=i == 0)
        _check_row(r, row)
        if i == 0:
            for name in FIELD_NAMES + ["flux"]:
                assert np.array_equal(r[name], np.asarray(row["fields"][name])), name


def test_symmetry_patch():
    """A symmetry plane (BoundaryCondition.hpp:279-384 symmetry branches)."""
    d = load("ref_channel_symmetry.json")
    prob = symmetry_channel()
    r = _oracle(prob).evaluate(prob.coefficients(d["theta"]), fields=True)
    _check_row(r, d["row"])
    for name in FIELD_NAMES + ["flux"]:
        assert np.array_equal(r[name], np.asarray(d["row"]["fields"][name])), name


def test_moving_wall_couette():
    """Moving top wall (WallMoving, couetteDefaults) on the graded 4x96
    channel with PCG pressure: rows and fields equal the reference's. The
    two walls carry the same shear in Couette flow, so the mean wall shear
    stresses of bottom and top agree."""
    d = load("ref_couette.json")
    prob = couette_problem()
    P = _oracle(prob)
    for i, (theta, row) in enumerate(zip(d["thetas"], d["rows"])):
        r = P.evaluate(prob.coefficients(theta), fields=i == 0)
        _check_row(r, row)
        if i == 0:
            for name in FIELD_NAMES + ["flux"]:
                assert np.array_equal(r[name], np.asarray(row["fields"][name])), name


def test_free_stream_flat_plate():
    """A Dirichlet velocity, k and omega on a patch that is not an inlet (the
    free-stream top of flatPlateDefaults) on the graded plate mesh: rows and
    fields equal the reference's, and the top faces carry the free stream."""
    d = load("ref_plate.json")
    prob = plate_problem()
    P = _oracle(prob)
    for i, (theta, row) in enumerate(zip(d["thetas"], d["rows"])):
        r = P.evaluate(prob.coefficients(theta), fields=i == 0)
        _check_row(r, row)
        if i == 0:
            for name in FIELD_NAMES + ["flux"]:
                assert np.array_equal(r[name], np.asarray(row["fields"][name])), name
    f = d["rows"][0]["fields"]
    assert len(f["Uxb"]) == prob.mesh.n_faces - prob.mesh.n_internal
    top = next(p for p in prob.mesh.patches if p.name == "top")
    b = np.asarray(top.faces) - prob.mesh.n_internal
    assert len(b) == 24
    assert np.all(np.asarray(f["Uxb"])[b] == d["u_inf"]) and np.all(np.asarray(f["Uyb"])[b] == 0)
    assert np.all(np.asarray(f["kb"])[b] == d["k_in"])
    assert np.all(np.asarray(f["omegab"])[b] == d["omega_in"])


def test_backward_facing_step():
    """The reference's backward-facing-step mesh (step face, re-entrant
    corner, symmetry top, inlet and outlet), 600 capped iterations: the row
    and every field equal the reference's."""
    d = load("ref_bfs.json")
    prob = bfs_problem()
    assert {p.name: p.type for p in prob.mesh.patches}["top_wall"] == "symmetry"
    r = _oracle(prob).evaluate(prob.coefficients(d["theta"]), fields=True)
    _check_row(r, d["row"])
    for name in FIELD_NAMES + ["flux"]:
        assert np.array_equal(r[name], np.asarray(d["row"]["fields"][name])), name


@pytest.mark.parametrize("case", ["cosine_bump", "bfs"])
def test_separation_and_reattachment(case):
    """Separation-point and reattachment-length observations on the cosine
    bump and the backward-facing step: rows (these observations included)
    equal the reference's; behind the step the bubble reattaches inside the
    domain, and the wall before the step never separates (the value 0)."""
    d = load("ref_wall_crossings.json")[case]
    prob = crossings_problems()[case]
    P = _oracle(prob)
    for i, (theta, row) in enumerate(zip(d["thetas"], d["rows"])):
        r = P.evaluate(prob.coefficients(theta), fields=i == 0)
        _check_row(r, row)
    if case == "bfs":
        # a recirculation bubble behind the step, reattaching within the
        # domain, and no separation on the wall before the step
        x_r = r["predictions"][0]
        assert 1.0 < x_r < 22.0 and r["predictions"][2] == 0.0


@pytest.mark.parametrize("case", ["plate", "bfs"])
def test_drag(case):
    """Drag observations on the flat plate and on the step face and lower
    wall of the backward-facing step: rows equal the reference's. On the
    plate the pressure term vanishes (the wall normal has no x part), so the
    drag is the area-weighted mean skin friction of the converged field."""
    d = load("ref_drag.json")[case]
    prob = drag_problems()[case]
    P = _oracle(prob)
    for i, (theta, row) in enumerate(zip(d["thetas"], d["rows"])):
        r = P.evaluate(prob.coefficients(theta), fields=i == 0)
        _check_row(r, row)
        if case == "plate" and i == 0:
            m = prob.mesh
            wall = next(p for p in m.patches if p.name == "bottom").faces
            ow = m.owner[wall]
            assert np.all(m.face_normal[wall][:, 0] == 0.0)
            # a stationary no-slip wall: the boundary velocity is zero
            tau = prob.nu * np.asarray(r["Ux"])[ow] / np.maximum(m.face_delta[wall], 1e-20)
            cd = float(np.sum(tau * m.face_area[wall])) / 0.5
            assert r["predictions"][0] == pytest.approx(cd, rel=1e-12)


def test_reference_couette_golden():
    """The reference's committed Couette regression (couette_re171_ny96, its
    AMG pressure solver, recorded with another compiler and libm): status,
    iteration count and stalled flag exactly, values to the last few digits,
    and the wall shear balance it reports (tau_w bottom = top to 1e-3)."""
    g = load("ref_couette_re171_ny96.json")
    prob = couette_problem(pressure_solver="AMG")
    assert prob.nu == g["nu"] and prob.mesh.n_cells == 4 * g["mesh"]["ny"]
    r = _oracle(prob).evaluate(prob.coefficients([0.31, 0.09]), fields=True)
    res = g["result"]
    assert r["status"] == 0 and r["iters"] == res["final_iter"]
    assert r["stalled"] == res["stalled"]
    q = g["qoi"]
    assert max_rel(r["predictions"], [q["cf_probe"]] + q["u_probes"]) < 1e-9
    assert r["worst"] == pytest.approx(res["final_worst_measure"], rel=1e-5)
    assert r["rate"] == pytest.approx(res["contraction_rate_u"], rel=1e-9)
    norms = res["field_norms"]
    assert float(np.sum(r["Ux"] ** 2)) == pytest.approx(norms["sumsq_ux"], rel=1e-9)
    assert float(np.sum(r["k"] ** 2)) == pytest.approx(norms["sumsq_k"], rel=1e-9)
    assert float(np.sum(r["omega"] ** 2)) == pytest.approx(norms["sumsq_omega"], rel=1e-9)
    assert float(np.sum(r["nuT"] ** 2)) == pytest.approx(norms["sumsq_nut"], rel=1e-9)
    assert q["tau_w_top"] == pytest.approx(q["tau_w_bottom"], rel=1e-3)


def test_wall_normal_body_force():
    """A y body force on the periodic channel and on the developing channel
    (the momentum source of SIMPLESolver.cpp:303-482): every row and field
    equals the reference's, and on the periodic channel the force is held by
    the pressure: away from the walls p rises by f_y per unit height."""
    d = load("ref_channel_body_force.json")
    c = d["periodic"]
    prob = small_channel()
    fy = c["body_force"][1]
    prob.settings.body_force.y = fy
    r = _oracle(prob).evaluate(prob.coefficients(c["theta"]), fields=True)
    _check_row(r, c["row"])
    for name in FIELD_NAMES + ["flux"]:
        assert np.array_equal(r[name], np.asarray(c["row"]["fields"][name])), name
    y = prob.mesh.cell_center[:, 1]
    order = np.argsort(y)
    dpdy = np.diff(np.asarray(r["p"])[order]) / np.diff(y[order])
    core = dpdy[len(dpdy) // 4: -(len(dpdy) // 4)]
    assert np.all(np.abs(core - fy) <= 5e-4 * fy), core
    io = d["inout"]
    prob = inout_channel("upwind", "upwind")
    prob.settings.body_force.y = io["body_force"][1]
    P = _oracle(prob)
    for i, (theta, row) in enumerate(zip(io["thetas"], io["rows"])):
        r = P.evaluate(prob.coefficients(theta), fields=i == 0)
        _check_row(r, row)
        if i == 0:
            for name in FIELD_NAMES + ["flux"]:
                assert np.array_equal(r[name], np.asarray(row["fields"][name])), name


def test_converged_small_hill_amg():
    """Converged 30x16 hill with the reference's AMG pressure solver."""
    d = load("ref_hill_small_converged.json")
    prob = cases.hill_problem(nx=30, ny=16)
    _check_row(_oracle(prob).evaluate(prob.coefficients(d["thetas"][0])), d["rows"][0])


def test_channel_draws_fixture():
    d = load("channel_theta64.json")
    assert d["seed"] == cases.CHANNEL_SEED
    assert np.array_equal(cases.channel_draws(64), np.asarray(d["theta"]))


# ----------------------------------------------------------- physics invariants
def test_channel_wall_shear_balances_body_force():
    """Integrated momentum balance of the converged periodic channel: the two
    walls carry the body force, tau_w = f h = 1 (test_periodic_channel_retau.cpp:83-97),
    to within the discretisation of the first-cell gradient."""
    prob = small_channel()
    r = _oracle(prob).evaluate(prob.coefficients([0.31, 0.09]), fields=True)
    m = prob.mesh
    wall = np.arange(m.n_internal, m.n_faces)
    tau = []
    for f in wall:
        c = m.owner[f]
        tau.append(prob.nu * r["Ux"][c] / m.face_delta[f])
    tau = np.asarray(tau)
    assert np.all(np.abs(tau - 1.0) < 0.05)
    # wall shear of the two walls is symmetric
    assert abs(tau[0] - tau[-1]) < 1e-8


def test_one_column_equals_three_columns():
    """The one-column channel (self-wrapped periodic faces) reaches the same
    fixed point as the three-column mesh of the same wall-normal grading. The
    iteration counts differ (round-off cross-stream pressure noise on three
    columns keeps the y-momentum and p' solves busy), the probes agree to the
    solve tolerance (at ny=128, tol 1e-8 they agree to 3e-9)."""
    p1 = cases.channel_problem(ny=32, nx=1, ratio=1.15, n_probes=30, tol=1e-6, max_iterations=40000)
    p3 = cases.channel_problem(ny=32, nx=3, ratio=1.15, n_probes=30, tol=1e-6, max_iterations=40000)
    r1 = _oracle(p1).evaluate(p1.coefficients([0.31, 0.09]), fields=True)
    r3 = _oracle(p3).evaluate(p3.coefficients([0.31, 0.09]), fields=True)
    assert r1["status"] == r3["status"] == 0
    assert max_rel(r3["predictions"], r1["predictions"]) < 1e-6
    # streamwise invariance of the three-column solution
    ux = r3["Ux"].reshape(32, 3) if p3.mesh.n_cells == 96 else None
    assert ux is not None
    assert np.max(np.abs(ux - ux[:, :1])) <= 1e-8 * np.max(np.abs(ux))
