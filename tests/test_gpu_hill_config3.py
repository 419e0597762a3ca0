"""The 16-draw periodic hill at its full size (120x64 cells, AMG pressure,
tol 1e-4, cap 25,000) against the reference's own solver on the same draws
(ref_hill_c3_converged.json, written by tests/golden/make_fixtures.py
hill_c3; the reference's hills case, tests/tools/record_goldens.cpp:558-628).

The bar:
* every status equals the reference's (10 Converged, 6 Unconverged at the
  cap), in FP64 and with FP32 Krylov plus FP64 defect correction;
* converged rows: the outer iteration count within max(20, 0.5 %) (a
  monitor interval; the engine's parallel reductions round differently from
  the reference's sequential sums) and every probe within 1e-6 relative;
* rows at the cap: 25,000 iterations and the stalled flag exactly. Their
  probes are the state of an unfinished iteration, which depends on the
  iteration path, and the device's AMG smooths with damped Jacobi where the
  reference smooths with Gauss-Seidel, so each p' solve stops at the same
  relative residual (1e-4) with a different iterate. Those probes are held
  to 5e-3 (measured up to 1.3e-3 in both precisions); such rows are never
  admitted to a posterior.
"""
from __future__ import annotations

import numpy as np
import pytest

from cfdr_amd import cases, engine
from fixtures import load, max_rel

pytestmark = pytest.mark.gpu

REL_OBS = 1e-6
REL_CAPPED = 5e-3


def _check(mixed):
    d = load("ref_hill_c3_converged.json")
    prob = cases.hill_problem(nx=120, ny=64)
    b = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                            prob.u_init, prob.p_init, prob.k_init, prob.omega_init,
                            mixed_precision=mixed)
    res = b.evaluate(np.asarray(d["thetas"]))
    b.close()
    worst_converged = 0.0
    for i, row in enumerate(d["rows"]):
        got = int(res.statuses[i])
        print(f"row {i}: status {got}/{row['status']} iterations "
              f"{int(res.iteration_counts[i])}/{row['iters']} probes "
              f"{max_rel(res.predictions[i], row['predictions']):.2e}", flush=True)
        assert got == row["status"], i
        if row["status"] == 0:
            assert abs(int(res.iteration_counts[i]) - row["iters"]) <= max(20, int(0.005 * row["iters"]))
            worst_converged = max(worst_converged, max_rel(res.predictions[i], row["predictions"]))
        else:
            assert int(res.iteration_counts[i]) == row["iters"] == prob.settings.max_iterations
            assert bool(res.stalled[i]) == row["stalled"], i
            assert max_rel(res.predictions[i], row["predictions"]) <= REL_CAPPED, i
    assert worst_converged <= REL_OBS


def test_config3_hill_fp64_matches_reference():
    _check(False)


def test_config3_hill_mixed_precision_matches_reference():
    _check(True)
