"""Why the assembly stays FP64: the converged/failed decision of
SIMPLESolver::solve (rans-uq/src/SIMPLESolver.cpp:1545-1705) rests on the
contraction of successive field changes (ContractionEstimate,
SIMPLESolver.cpp:32-74). Rounding every assembled system to FP32, which is
the least an FP32 assembly costs (its coefficients and sources are stored
in FP32 whatever precision forms them), puts a floor under those changes:
the estimated contraction rate reaches 1 and the draw never converges,
whatever the tolerance, although its probes stay within 1e-6 of the FP64
solve. The oracle carries the rounding as an experiment switch
(or_set_assembly_rounding); everything else, residuals and norms included,
stays FP64 in both runs.

Measured on the first three draws of the 64-draw sweep at tol 1e-8: FP64
converges in 41,690 / 44,330 / 45,150 outer iterations; with the rounding
every draw stalls at the 150,000 cap (rate 1.0001-1.005) with probes
5e-7-8e-7 from FP64. At tol 1e-6, 1e-7 and 3e-8 the rounded run also
stalls (at tol 1e-6: rate 0.998 after 55,800 iterations, 1.004 after
80,000). The test below repeats the 1e-6 case with an 80,000 cap.
"""
from __future__ import annotations

import numpy as np

import oracle_py
from cfdr_amd import cases


def test_fp32_rounded_assembly_never_converges():
    prob = cases.channel_problem()
    o = oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs, prob.nu,
                                prob.init_vector())
    c = prob.coefficients(cases.channel_draws(64)[0])
    L = oracle_py.lib()
    ref = o.evaluate(c, tol=1e-6)
    assert ref["status"] == 0 and ref["iters"] == 27900
    try:
        L.or_set_assembly_rounding(1)
        got = o.evaluate(c, tol=1e-6, max_iter=80000)
    finally:
        L.or_set_assembly_rounding(0)
    # the run stops at the cap, not converged, with its contraction rate at 1
    assert got["status"] == 1 and got["iters"] == 80000
    assert got["stalled"] and got["rate"] >= 1.0
    rel = np.max(np.abs(got["predictions"] - ref["predictions"]) / np.abs(ref["predictions"]))
    assert rel < 1e-5
    # the switch is off again: the FP64 result comes back bit for bit
    again = o.evaluate(c, tol=1e-6)
    assert again["iters"] == ref["iters"]
    assert np.array_equal(again["predictions"], ref["predictions"])
