"""Host-side logic and the C-ABI boundary, without a GPU.

* the engine library loads and exports every function include/rans_amd.h
  declares (no compute call is made here);
* argument validation that the engine performs before touching a device
  (the reference's std::invalid_argument cases of ForwardBatch.hpp:53-58 and
  SolverSettings::validate, FlowFields.hpp:203-240);
* the host package: MSH2 round trip, parameter sets, observation stencils
  (linear-field exactness, core/tests/test_probe_interpolation.cpp:139) and
  the log-likelihood conventions (ObservationOperator.hpp:592-607).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from cfdr_amd import cases, engine
from cfdr_amd.config import (COEFF_DEFAULTS, InferenceParameterSet, ObservationOperator,
                             SolverSettings, Vec3)
from cfdr_amd.mesh import Mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rans_amd.h")


# ------------------------------------------------------------------ C-ABI
def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rq_[a-z_]+)\s*\(", text)))


def test_header_declares_exactly_the_binding_exports():
    assert _declared_functions() == sorted(engine.EXPORTS)


def test_library_loads_and_exports_every_symbol():
    lib = engine.load_library()
    for name in _declared_functions():
        assert hasattr(lib, name), name
    info = engine.build_info()
    assert "gfx950" in info


def test_library_is_not_linked_to_the_oracle():
    with open(engine.LIB_PATH, "rb") as fh:
        blob = fh.read()
    assert b"liboracle" not in blob and b"or_evaluate" not in blob


def _raw_create(settings_mut=None, theta_mut=None):
    """Calls rq_engine_create with a small channel whose settings are mutated
    to an invalid value; validation fails before any device call."""
    prob = cases.channel_problem(ny=8, ratio=1.2, n_probes=4, tol=1e-6, max_iterations=10)
    keep = engine._Keep()
    m = engine.mesh_struct(prob.mesh, keep)
    b = engine.RqBc()
    ut, uv, pt, pv, kt, kv, wt, wv = prob.bcs.arrays()
    b.u_type, b.p_type, b.k_type, b.w_type = keep.i(ut), keep.i(pt), keep.i(kt), keep.i(wt)
    b.u_val, b.p_val, b.k_val, b.w_val = keep.d(uv), keep.d(pv), keep.d(kv), keep.d(wv)
    p = engine.RqParams()
    ps = prob.params
    p.n_active = ps.n_active()
    p.active_indices = keep.i(ps.active_indices)
    p.base = keep.d(ps.base)
    p.lower, p.upper = keep.d(ps.lower_bounds()), keep.d(ps.upper_bounds())
    o = engine.obs_struct(prob.obs.prepare(prob.mesh), keep)
    s = engine.settings_struct(prob.settings)
    if settings_mut:
        settings_mut(s)
    init = np.asarray(prob.init_vector(), float)
    err = C.create_string_buffer(512)
    lib = engine.load_library()
    h = lib.rq_engine_create(C.byref(m), C.byref(p), C.byref(o), C.byref(b), prob.nu, C.byref(s),
                             init.ctypes.data_as(engine._dp), err, 512)
    return h, err.value.decode()


@pytest.mark.parametrize("field,value,text", [
    ("convergence_window", 1, "iteration count or cadence"),
    ("divergence_limit", -1.0, "divergence limit"),
    ("alpha_u", 0.0, "relaxation denominators"),
    ("k_max", 1e-12, "kMax"),
    ("momentum_solver", 2, "pressure correction only"),
    ("block_size", 96, "block_size"),
    ("convergence_window", 300, "engine limit"),
    ("amg_block_rows", -1, "amg_block_rows"),
    ("grid_blocks", 65, "grid_blocks"),
])
def test_create_rejects_invalid_settings(field, value, text):
    def mut(s):
        setattr(s, field, value)
    h, msg = _raw_create(mut)
    assert not h
    assert text in msg


def test_python_front_end_validation():
    prob = cases.channel_problem(ny=8, ratio=1.2, n_probes=4, tol=1e-6, max_iterations=10)
    s = SolverSettings(momentum_solver="AMG")
    with pytest.raises(ValueError, match="pressure correction only"):
        engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, s)
    with pytest.raises(ValueError, match="viscosity"):
        engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, -1.0, prob.settings)
    m = prob.mesh.copy()
    m.wall_distance_ = None
    if not m.wall_distance_valid():
        with pytest.raises(ValueError, match="compute_wall_distance"):
            engine.ForwardBatch(m, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings)


# ------------------------------------------------------------ host package
def test_msh2_round_trip(tmp_path):
    m = Mesh.make_periodic_channel_2d(3, 10, 1.0, 1.0, 182.088, 0.5)
    path = str(tmp_path / "m.msh")
    m.save_binary(path)
    r = Mesh.load_binary(path)
    for a in ("owner", "neighbor", "face_center", "face_normal", "face_area", "face_delta",
              "face_d", "face_weight", "face_shift", "cell_center", "cell_volume",
              "cell_face_ptr", "cell_face_idx"):
        assert np.array_equal(getattr(m, a), getattr(r, a)), a
    assert [(p.name, p.type, list(p.faces)) for p in m.patches] == \
        [(p.name, p.type, list(p.faces)) for p in r.patches]
    with open(path, "ab") as fh:
        fh.write(b"\0")
    with pytest.raises(ValueError, match="trailing"):
        Mesh.load_binary(path)


def test_geometric_channel_grading():
    m = Mesh.make_geometric_channel_2d(1, 128, 1.0, 1.0, 1.05)
    y = m.cell_center[:, 1]
    assert m.n_cells == 128 and np.all(np.diff(y) > 0)
    horizontal = np.abs(m.face_normal[:, 1]) == 1.0
    dy = np.diff(np.unique(m.face_center[horizontal, 1]))
    assert len(dy) == 128
    assert dy[1] / dy[0] == pytest.approx(1.05, rel=1e-9)
    assert dy[0] == pytest.approx(0.05 / (1.05 ** 64 - 1), rel=1e-9)
    assert np.allclose(y, 2.0 - y[::-1], atol=1e-14)


def test_parameter_sets():
    ps = InferenceParameterSet.a1_betaStar()
    assert ps.active_indices == [9, 8]
    full = ps.unpack([0.3, 0.08])
    assert full[9] == 0.3 and full[8] == 0.08 and full[0] == COEFF_DEFAULTS[0]
    assert ps.in_bounds([0.3, 0.08]) and not ps.in_bounds([0.9, 0.08])
    assert not ps.in_bounds([float("nan"), 0.08])
    with pytest.raises(ValueError, match="duplicate"):
        InferenceParameterSet.from_indices("x", [1, 1])
    with pytest.raises(ValueError, match="size mismatch"):
        ps.unpack([0.3])


def test_velocity_probe_reproduces_linear_field():
    """The least-squares probe stencil is exact for a linear field
    (core/tests/test_probe_interpolation.cpp:139). On a streamwise-periodic
    mesh the field is linear in y only, so it is periodic too; cell values of
    periodic neighbours and wall-face values enter as stored."""
    m = Mesh.make_periodic_channel_2d(4, 12, 1.0, 1.0, 182.088, 0.5)
    m.compute_wall_distance()
    ob = ObservationOperator()
    locs = [(0.37, 0.05), (0.5, 0.61), (0.9, 1.93), (0.12, 1.2), (0.01, 0.002)]
    for x, y in locs:
        ob.add_velocity_profile((x, y), 0, 0.0, 0.1)
    P = ob.prepare(m)

    def u(y):
        return 2.0 - 1.5 * y

    cu = u(m.cell_center[:, 1])
    fu = u(m.face_center[:, 1])
    for i, (x, y) in enumerate(locs):
        center = cu[P.cell[i]]
        bx = by = 0.0
        for q in range(P.nb_ptr[i], P.nb_ptr[i + 1]):
            nb = P.nb_cell[q]
            v = fu[P.nb_face[q]] if nb < 0 else cu[nb]
            bx += P.nb_w[q] * P.nb_dx[q] * (v - center)
            by += P.nb_w[q] * P.nb_dy[q] * (v - center)
        gx = (P.ayy[i] * bx - P.axy[i] * by) / P.det[i]
        gy = (P.axx[i] * by - P.axy[i] * bx) / P.det[i]
        assert abs(gx) < 1e-10 and gy == pytest.approx(-1.5, rel=1e-10)
        value = center + gx * P.offx[i] + gy * P.offy[i]
        assert value == pytest.approx(u(y), rel=1e-12)


def test_log_likelihood_conventions():
    ob = ObservationOperator()
    ob.add_velocity_profile((0.5, 0.5), 0, 1.0, 0.5, 0.5)
    assert ob.log_likelihood([2.0]) == pytest.approx(-0.5 * 1.0 / 0.5)
    with pytest.raises(ValueError):
        ob.log_likelihood([float("nan")])
    big = ObservationOperator()
    big.add_velocity_profile((0.5, 0.5), 0, 1e308, 1e-3)
    with pytest.raises(OverflowError):
        big.log_likelihood([-1e308])


def test_channel_problem_matches_the_single_solve_workload():
    p = cases.channel_problem()
    assert p.mesh.n_cells == 128 and p.settings.convergence_tol == 1e-8
    assert p.settings.pressure_solver == "PCG" and p.settings.momentum_solver == "BiCGSTAB"
    assert p.obs.n_obs() == 30 and p.params.active_indices == [9, 8]
    yplus = [ob.location[1] * cases.RE_TAU for ob in p.obs.observables]
    assert yplus[0] == pytest.approx(1.0) and yplus[-1] == pytest.approx(182.0)
    assert p.settings.body_force.x == 1.0 and p.settings.alpha_u == 0.5
    assert math.isclose(p.nu, 1.0 / 182.088)
    assert (p.u_init.x, p.k_init, p.omega_init) == (15.0, 1.0, 10.0)


def test_engine_extensions_reject_null_engines():
    """The autotuner and graph-replay entry points fail with RQ_EINVAL on a
    null engine, before any device call."""
    lib = engine.load_library()
    err = C.create_string_buffer(256)
    theta = np.zeros(2)
    out = engine.RqTuneOut()
    assert lib.rq_engine_autotune(None, theta.ctypes.data_as(engine._dp), 1, 2, 10, C.byref(out),
                                  err, 256) == 1
    assert lib.rq_engine_set_graph_replay(None, 1, err, 256) == 1
    assert "null engine" in err.value.decode()
    assert lib.rq_engine_graph_launches(None) == -1


def test_patch_observables_validation_and_stencils():
    """addDrag's reference checks (ObservationOperator.hpp:189-201), an
    unknown patch name, and the prepared stencil of the patch observables:
    the patch's faces in patch order with their centre x, drag's reference
    area in offx."""
    with pytest.raises(ValueError, match="drag reference area"):
        ObservationOperator().add_drag("bottom", 0.0, 0.1, 0.0, 1.0)
    with pytest.raises(ValueError, match="drag reference speed"):
        ObservationOperator().add_drag("bottom", 0.0, 0.1, 1.0, math.nan)
    with pytest.raises(ValueError, match="Gaussian measurements"):
        ObservationOperator().add_separation_point("bottom", 0.0, 0.0)
    prob = cases.channel_problem(ny=8, ratio=1.2, n_probes=1)
    mesh = prob.mesh
    op = ObservationOperator()
    op.add_reattachment_length("no_such_patch", 0.0, 0.1)
    with pytest.raises(ValueError, match="unknown patch"):
        op.prepare(mesh)
    op = ObservationOperator()
    op.add_drag("bottom", 0.0, 0.1, 2.5, 1.0)
    op.add_separation_point("top", 0.0, 0.1)
    P = op.prepare(mesh)
    assert list(P.type) == [3, 4] and P.offx[0] == 2.5
    for i, name in enumerate(("bottom", "top")):
        faces = next(p for p in mesh.patches if p.name == name).faces
        q = slice(P.nb_ptr[i], P.nb_ptr[i + 1])
        assert list(P.nb_face[q]) == list(faces) and np.all(P.nb_cell[q] == -1)
        assert np.array_equal(P.nb_dx[q], mesh.face_center[faces][:, 0])
