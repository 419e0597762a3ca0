"""The serving front end (rq_server_*) on the GPU: requests of uneven mesh
sizes admitted while others run, each result the same alone and in the mix,
parity with the reference fixtures, host-settled invalid rows, and the device
loop's own idle exit followed by a relaunch."""
from __future__ import annotations

import time

import numpy as np
import pytest

from cfdr_amd import cases, engine
from cfdr_amd.serve import Server
from fixtures import inout_channel, load, max_rel, small_channel

pytestmark = pytest.mark.gpu

REL_OBS = 1e-6


def _add(server, prob):
    return server.add_problem(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                              prob.u_init, prob.p_init, prob.k_init, prob.omega_init)


def _same(a, b):
    return (a.status == b.status and a.simple_iters == b.simple_iters
            and np.array_equal(a.predictions, b.predictions, equal_nan=True)
            and a.log_lik == b.log_lik)


def test_result_same_alone_and_in_mix():
    small = small_channel()
    three = cases.channel_problem(ny=32, nx=3, ratio=1.15, n_probes=30, tol=1e-6,
                                  max_iterations=40000)
    inout = inout_channel("upwind", "upwind")
    theta = cases.channel_draws(6)
    with Server(max_slots=64, idle_exit_seconds=1.0) as s:
        pids = [_add(s, p) for p in (small, three, inout)]
        alone = s.wait(s.submit(pids[2], theta[0]))
        alone_small = s.wait(s.submit(pids[0], theta[1]))
        mix = [(0, 2), (1, 3), (2, 0), (0, 4), (1, 5), (0, 1), (2, 3), (1, 0), (0, 5), (2, 0),
               (1, 2), (0, 1)]
        tickets = []
        for which, k in mix:
            tickets.append((pids[which], theta[k], s.submit(pids[which], theta[k])))
        mixed = [(pid, th, s.wait(t)) for pid, th, t in tickets]
        stats = s.stats()
    hits = 0
    for pid, th, r in mixed:
        if pid == pids[2] and np.array_equal(th, theta[0]):
            assert _same(r, alone)
            hits += 1
        if pid == pids[0] and np.array_equal(th, theta[1]):
            assert _same(r, alone_small)
            hits += 1
    assert hits == 4
    assert stats["completed"] == 14 and stats["slots"] >= 1


def test_served_rows_match_reference():
    d = load("ref_channel_small.json")
    with Server(max_slots=16, idle_exit_seconds=1.0) as s:
        pid = _add(s, small_channel())
        tickets = [s.submit(pid, th) for th in d["thetas"]]
        rows = [s.wait(t) for t in tickets]
    for r, row in zip(rows, d["rows"]):
        assert r.status == row["status"]
        assert abs(r.simple_iters - row["iters"]) <= max(20, int(0.005 * row["iters"]))
        assert max_rel(r.predictions, row["predictions"]) <= REL_OBS
        assert r.log_lik == pytest.approx(row["loglik"], rel=REL_OBS)


def test_served_matches_batch_path():
    prob = small_channel()
    theta = cases.channel_draws(4)
    batch = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                                prob.u_init, prob.p_init, prob.k_init, prob.omega_init)
    ref = batch.evaluate(theta)
    with Server(max_slots=8, idle_exit_seconds=1.0) as s:
        pid = _add(s, prob)
        rows = [s.wait(s.submit(pid, th)) for th in theta]
    for i, r in enumerate(rows):
        assert r.status == ref.statuses[i]
        assert max_rel(r.predictions, ref.predictions[i]) <= REL_OBS


def test_invalid_rows_settle_on_host_and_idle_relaunch():
    prob = small_channel()
    with Server(max_slots=4, idle_exit_seconds=0.2) as s:
        pid = _add(s, prob)
        bad = s.poll(s.submit(pid, [0.9, 0.09]))
        assert bad is not None and bad.status == engine.EvaluationStatus.InvalidParameters
        assert bad.log_lik == -1e30 and np.all(np.isnan(bad.predictions))
        first = s.wait(s.submit(pid, [0.31, 0.09]))
        time.sleep(1.0)  # the device loop has returned on its idle limit
        second = s.wait(s.submit(pid, [0.31, 0.09]))
        assert _same(first, second)
        with pytest.raises(ValueError):
            s.submit(pid, [0.31])
        with pytest.raises(ValueError, match="before the first submit"):
            _add(s, prob)


def test_served_hill_in_a_channel_mix_matches_batch_path():
    """A 60x32 periodic hill (AMG pressure, arena and multigrid workspace in
    HBM; solves capped at 120 outer iterations) served among channel
    requests: each hill row equals the one-workgroup batch path's row."""
    hill = cases.hill_problem(nx=60, ny=32, max_iterations=120)
    small = small_channel()
    th_h = cases.hill_draws(3)
    th_c = cases.channel_draws(3)
    batch = engine.ForwardBatch(hill.mesh, hill.params, hill.obs, hill.bcs, hill.nu,
                                hill.settings, hill.u_init, hill.p_init, hill.k_init,
                                hill.omega_init, block_size=256, grid_blocks=1)
    ref = batch.evaluate(th_h)
    with Server(max_slots=16, idle_exit_seconds=1.0) as s:
        ph, pc = _add(s, hill), _add(s, small)
        tickets = []
        for i in range(3):
            tickets.append(("h", i, s.submit(ph, th_h[i])))
            tickets.append(("c", i, s.submit(pc, th_c[i])))
        rows = [(kind, i, s.wait(t)) for kind, i, t in tickets]
    # the hill makes the loop run 512-thread workgroups, where the 32-cell
    # channel takes the block Krylov path instead of the one-wave path
    cref = engine.ForwardBatch(small.mesh, small.params, small.obs, small.bcs, small.nu,
                               small.settings, small.u_init, small.p_init, small.k_init,
                               small.omega_init).evaluate(th_c)
    for kind, i, r in rows:
        if kind == "c":
            assert r.status == cref.statuses[i]
            assert max_rel(r.predictions, cref.predictions[i]) <= REL_OBS
            continue
        assert r.status == ref.statuses[i] == engine.EvaluationStatus.Unconverged
        assert r.simple_iters == ref.iteration_counts[i] == 120
        assert max_rel(r.predictions, ref.predictions[i]) <= REL_OBS
