"""HIP graph replay of the batch launch (rq_engine_set_graph_replay, no
reference counterpart): every replay gives the bits of a directly launched
evaluation of the same rows, with other GPU work and host synchronisation
between the replays; a changed row count captures a new graph; draws spread
over several workgroups keep the direct launch."""
from __future__ import annotations

import numpy as np
import pytest

from cfdr_amd import cases, engine
from fixtures import small_channel, small_hill

pytestmark = pytest.mark.gpu


def _batch(prob, **kw):
    return engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                               prob.u_init, prob.p_init, prob.k_init, prob.omega_init, **kw)


def _same(a, b):
    assert [int(s) for s in a.statuses] == [int(s) for s in b.statuses]
    assert np.array_equal(a.iteration_counts, b.iteration_counts)
    assert np.array_equal(a.predictions, b.predictions, equal_nan=True)
    assert np.array_equal(a.log_likelihoods, b.log_likelihoods, equal_nan=True)
    assert np.array_equal(a.krylov_iterations, b.krylov_iterations)


def test_replays_match_direct_launches():
    prob = small_channel()
    direct = _batch(prob)
    g = _batch(prob)
    g.set_graph_replay(True)
    draws = cases.channel_draws(12)
    sets = [draws[0:4], draws[4:8], draws[8:12], draws[0:4]]
    for k, th in enumerate(sets):
        got = g.evaluate(th, fields=True)
        assert g.graph_launches() == k + 1
        want = direct.evaluate(th, fields=True)  # other GPU work between replays
        _same(got, want)
        assert np.array_equal(got.fields, want.fields)
    # a new row count captures a new graph; its replays still match
    for th in (draws[0:3], draws[3:6]):
        _same(g.evaluate(th), direct.evaluate(th))
    assert g.graph_launches() == 6
    # rows rejected on the host never reach the graph
    bad = np.vstack([draws[0], [1e9, 0.09], draws[1]])
    got, want = g.evaluate(bad), direct.evaluate(bad)
    _same(got, want)
    assert int(got.statuses[1]) == int(engine.EvaluationStatus.InvalidParameters)


def test_multi_workgroup_draws_launch_directly():
    prob = small_hill("AMG")
    prob.settings.max_iterations = 50
    th = np.asarray([[0.31, 0.09, 0.85, 0.5], [0.3, 0.095, 0.9, 0.55]])
    g = _batch(prob, grid_blocks=4)
    g.set_graph_replay(True)
    a = g.evaluate(th)
    assert g.last_timing()["grid_blocks"] == 4 and g.graph_launches() == 0
    _same(a, _batch(prob, grid_blocks=4).evaluate(th))
    one = _batch(prob, grid_blocks=1)
    one.set_graph_replay(True)
    _same(one.evaluate(th), _batch(prob, grid_blocks=1).evaluate(th))
    assert one.graph_launches() == 1
    one.set_graph_replay(False)
    one.evaluate(th)
    assert one.graph_launches() == 1


def test_mixed_precision_replays_match_direct_launches():
    prob = small_channel()
    th = cases.channel_draws(4)
    g = _batch(prob, mixed_precision=True)
    g.set_graph_replay(True)
    direct = _batch(prob, mixed_precision=True)
    for _ in range(2):
        _same(g.evaluate(th), direct.evaluate(th))
    assert g.graph_launches() == 2


def test_concurrent_evaluates_on_one_engine():
    """Several Python threads evaluate on one engine (ctypes releases the
    GIL during the call): each result equals its own sequential evaluate."""
    import threading
    prob = small_channel()
    b = _batch(prob)
    draws = cases.channel_draws(8)
    sets = [draws[0:2], draws[2:4], draws[4:6], draws[6:8]]
    want = [b.evaluate(th) for th in sets]
    got = [None] * len(sets)

    def run(k):
        got[k] = b.evaluate(sets[k])

    threads = [threading.Thread(target=run, args=(k,)) for k in range(len(sets))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    for k in range(len(sets)):
        _same(got[k], want[k])
