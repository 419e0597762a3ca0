"""The OpenFOAM 2D case on the GPU engine (through the C-ABI) against the
reference's own solve of it (ref_foam_channel.json: the reference's
Mesh::loadFromFile of tests/golden/foam_channel_24x12, a graded 24x12 mesh
of trapezoids of depth 0.5 whose front and back faces form an "empty"
patch, solved with the developing-channel settings).

The engine folds the front and back faces into zero-gradient terms of the
momentum, k and omega rows (Mesh.fold_empty_faces; the CPU oracle does the
same and reproduces the reference's solve, tests/test_openfoam.py). The bar
is the parity bar of every other case: statuses exact, converged iteration
counts within max(20, 0.5 %), probes within 1e-6, the fields of a capped
40-iteration state within 1e-9 of each field's magnitude (the flux on the
in-plane faces, which come first in the case's face order). The folded
terms are compiled into the one-workgroup kernels only: a request for
several workgroups per draw runs on one, with the same bits.
"""
from __future__ import annotations

import numpy as np
import pytest

from cfdr_amd import engine
from cfdr_amd.serve import Server
from fixtures import FIELD_NAMES, foam_channel, load, max_rel, max_scaled

pytestmark = pytest.mark.gpu

REL_OBS = 1e-6
FIELD_TOL = 1e-9


def _batch(prob, **kw):
    return engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                               prob.u_init, prob.p_init, prob.k_init, prob.omega_init, **kw)


@pytest.mark.parametrize("mixed", [False, True])
def test_openfoam_case_rows_match_reference(mixed):
    d = load("ref_foam_channel.json")
    prob = foam_channel()
    b = _batch(prob, mixed_precision=mixed)
    res = b.evaluate(np.asarray(d["thetas"]))
    b.close()
    for i, row in enumerate(d["rows"]):
        print(f"mixed={mixed} row {i}: {res.statuses[i].name} {int(res.iteration_counts[i])}/"
              f"{row['iters']} probes {max_rel(res.predictions[i], row['predictions']):.2e}",
              flush=True)
        assert int(res.statuses[i]) == row["status"] == 0, res.errors[i]
        assert abs(int(res.iteration_counts[i]) - row["iters"]) <= max(20, int(0.005 * row["iters"]))
        assert max_rel(res.predictions[i], row["predictions"]) <= REL_OBS
        assert res.log_likelihoods[i] == pytest.approx(row["loglik"], rel=REL_OBS)


def test_openfoam_case_capped_fields_match_reference():
    d = load("ref_foam_channel.json")["capped"]
    prob = foam_channel(max_iterations=d["max_iterations"])
    b = _batch(prob)
    res = b.evaluate(np.asarray([d["theta"]]), fields=True)
    b.close()
    row = d["rows"][0]
    assert int(res.statuses[0]) == row["status"] and int(res.iteration_counts[0]) == row["iters"]
    nc = prob.mesh.n_cells
    f = res.fields[0]
    for j, name in enumerate(FIELD_NAMES):
        assert max_scaled(f[j * nc:(j + 1) * nc], row["fields"][name]) <= FIELD_TOL, name
    flux = f[10 * nc:]
    assert max_scaled(flux, row["fields"]["flux"][:flux.size]) <= FIELD_TOL


def test_openfoam_case_runs_one_workgroup_per_draw_with_the_same_bits():
    d = load("ref_foam_channel.json")
    prob = foam_channel()
    one = _batch(prob)
    r1 = one.evaluate(np.asarray(d["thetas"]))
    one.close()
    many = _batch(prob, grid_blocks=4)
    r4 = many.evaluate(np.asarray(d["thetas"]))
    assert many.last_timing()["grid_blocks"] == 1
    many.close()
    assert np.array_equal(r1.predictions, r4.predictions)
    assert list(r1.iteration_counts) == list(r4.iteration_counts)


def test_openfoam_case_served_equals_batch():
    d = load("ref_foam_channel.json")
    prob = foam_channel()
    b = _batch(prob)
    rb = b.evaluate(np.asarray(d["thetas"]))
    b.close()
    with Server(max_slots=8, idle_exit_seconds=1.0) as s:
        pid = s.add_problem(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                            prob.u_init, prob.p_init, prob.k_init, prob.omega_init)
        got = [s.wait(s.submit(pid, np.asarray(th))) for th in d["thetas"]]
    for i, r in enumerate(got):
        assert int(r.status) == int(rb.statuses[i])
        assert r.simple_iters == int(rb.iteration_counts[i])
        assert np.array_equal(r.predictions, rb.predictions[i])
