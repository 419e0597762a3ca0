"""The 256-thread LDS kernels assemble the momentum, k and omega rows with
two lanes per cell (each lane evaluates two face slots; the pair shares the
face terms with DPP broadcasts and adds them in slot order with the same
operations as the one-lane rows). A row's values are therefore bitwise the
same as with one lane per cell; only the order in which the block
reductions (norms, residual sums) meet differs with the launch shape, as
between any two block sizes.

The bar: the reference's fixtures at the parity bar with block_size 256
(statuses, iteration counts, probes within 1e-6, intermediate fields within
1e-9 of each field's magnitude), and the 64-draw sweep at 256 threads
against 128 threads (statuses, probes within 1e-10)."""
from __future__ import annotations

import numpy as np
import pytest

from cfdr_amd import cases, engine
from fixtures import FIELD_NAMES, load, max_rel, max_scaled, small_channel

pytestmark = pytest.mark.gpu

REL_OBS = 1e-6
FIELD_TOL = 1e-9


def _batch(prob, **kw):
    return engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                               prob.u_init, prob.p_init, prob.k_init, prob.omega_init, **kw)


def _iters_close(got, want):
    return abs(int(got) - int(want)) <= max(20, int(0.005 * want))


def test_small_channel_rows_and_states_match_reference_at_256_threads():
    d = load("ref_channel_small.json")
    prob = small_channel()
    b = _batch(prob, block_size=256)
    res = b.evaluate(np.asarray(d["thetas"]))
    t = b.last_timing()
    b.close()
    assert t["block_size"] == 256 and t["arena_in_lds"]
    for i, row in enumerate(d["rows"]):
        assert int(res.statuses[i]) == row["status"]
        if row["status"] == 0:
            assert _iters_close(res.iteration_counts[i], row["iters"])
        if row["status"] in (0, 1):
            assert max_rel(res.predictions[i], row["predictions"]) <= REL_OBS
    steps = load("ref_channel_steps.json")
    for K in ("2", "7", "600"):
        p = small_channel(max_iterations=int(K))
        bk = _batch(p, block_size=256)
        r = bk.evaluate(np.asarray([steps["theta"]]), fields=True)
        bk.close()
        nc = p.mesh.n_cells
        f = r.fields[0]
        for j, name in enumerate(FIELD_NAMES):
            assert max_scaled(f[j * nc:(j + 1) * nc], steps["steps"][K]["fields"][name]) \
                <= FIELD_TOL, (K, name)


def test_config1_channel_rows_match_reference_at_256_threads():
    d = load("ref_channel_c1.json")
    prob = cases.channel_problem()
    for mixed in (False, True):
        b = _batch(prob, block_size=256, mixed_precision=mixed)
        res = b.evaluate(np.asarray(d["thetas"]))
        b.close()
        for i, row in enumerate(d["rows"]):
            assert int(res.statuses[i]) == row["status"], (mixed, i)
            if row["status"] == 0:
                assert _iters_close(res.iteration_counts[i], row["iters"])
            assert max_rel(res.predictions[i], row["predictions"]) <= REL_OBS


def test_sweep_at_256_threads_matches_128_threads():
    prob = cases.channel_problem()
    theta = cases.channel_draws(64)
    a = _batch(prob, block_size=128)
    ra = a.evaluate(theta)
    a.close()
    b = _batch(prob, block_size=256)
    rb = b.evaluate(theta)
    b.close()
    assert [int(s) for s in rb.statuses] == [int(s) for s in ra.statuses]
    assert all(_iters_close(x, y) for x, y in zip(rb.iteration_counts, ra.iteration_counts))
    print("iteration counts differing:",
          int(np.sum(rb.iteration_counts != ra.iteration_counts)), "of 64; kernel ms",
          ra.kernel_ms, rb.kernel_ms, flush=True)
    assert max_rel(rb.predictions, ra.predictions) <= 1e-10


def test_three_column_channel_matches_oracle_at_256_threads():
    """Four faces per cell with distinct periodic neighbours: every slot of a
    row is a different face and a different neighbour."""
    import oracle_py
    prob = cases.channel_problem(ny=32, nx=3, ratio=1.15, n_probes=30, tol=1e-6,
                                 max_iterations=40000)
    theta = cases.channel_draws(3)
    b = _batch(prob, block_size=256)
    res = b.evaluate(theta)
    b.close()
    O = oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs, prob.nu,
                                prob.init_vector())
    for i in range(len(theta)):
        o = O.evaluate(prob.coefficients(theta[i]))
        assert int(res.statuses[i]) == o["status"]
        assert _iters_close(res.iteration_counts[i], o["iters"])
        assert max_rel(res.predictions[i], o["predictions"]) <= REL_OBS
