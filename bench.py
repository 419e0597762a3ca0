"""Throughput benchmark of the forward-solve engine on one MI355X.

A step is one batched evaluate of the 64-draw channel sweep (1x128
geometric channel, ratio 1.05, tol 1e-8, FP64): all 64 coefficient draws
run in one persistent-kernel launch. The line printed is

  value        converged solves/s of the batched path (whole job, one GPU)
  plain        the same draws solved one launch per draw (the plain per-solve
               GPU path), on a bounded sample of draws
  roofline     algorithmic bytes of the executed SIMPLE/Krylov stages per
               launch / kernel time (HIP events on the engine stream), against
               the 8 TB/s HBM peak; traffic is read from the committed
               rocprofv3 --pmc summary when one exists
  cpu_baseline the reference's own FP64 solver (compiled from its sources into
               oracle/_ref/ref_harness where it was built; otherwise the CPU
               FP64 oracle) on the host cores, a bounded sample of the same
               draws, one draw per process (thread); workers = the CPUs of the
               process's affinity mask within the job's CPU share
               (OMP_NUM_THREADS), with the host's counts in cpu_baseline.host;
               cpu_baseline_1worker times draw 0 alone on one process
  max_rel_err_vs_cpu_baseline  probe error of the GPU rows against the CPU rows
               of that sample (statuses must match)

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]

The sweep runs on the launch shape the autotuner picks (--no-autotune keeps the
default), and every step replays the HIP graph captured on the first
evaluate (--no-graph launches directly). Further legs (each can be switched
off): single_solve (one FP64 solve at the default coefficients, one launch,
with the oracle on one thread), mixed (the sweep with FP32 Krylov
and FP64 defect correction), served (Poisson arrivals of single requests on
64-512 cell channels against the plain one-launch-per-request path), hills
(the 16-draw 120x64 and 8-draw 480x256 hill batches per outer iteration,
and the 120x64 batch solved to convergence in FP64 and mixed precision,
whose converged solves/s the line repeats as hill_120x64_converged_solves_per_s;
--no-hills-full skips the latter), calibration (the staged calibration
driver on the 1x32 channel: solves, launches, seconds), warm_start (a
pattern-search poll of 4 neighbours of a converged point, cold in one batch
launch against warm-started from the device-resident state in one launch),
warm_fd_jacobian (the 22 perturbed solves of the eleven-coefficient warm-FD
Jacobian in one launch, against the oracle running them in sequence).

Profiling helpers: --max-iterations caps every solve (a short, fixed amount of
work for counter runs); --copy-calibration MIB runs only a device-to-device
copy of that size, whose FETCH_SIZE/WRITE_SIZE readings scale the counters of
the solve kernel to bytes.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bayesian-rans-uq_amd"))

HBM_PEAK_GBS = 8000.0
S_VAL, S_IDX = 8, 4  # FP64 values, int32 indices


# ------------------------------------------------------------ byte model
def _gg(N, F, B):
    """One Green-Gauss gradient: field on cells and boundary faces, face
    weight and S (3 doubles), owner/neighbour, gradient out (2 per cell)."""
    return (N + B) * S_VAL + F * 3 * S_VAL + 2 * F * S_IDX + 2 * N * S_VAL


def _assembly(N, F, k_in):
    """One equation's LDU assembly: face geometry (S, delta, w), owner and
    neighbour, k_in cell inputs, the face flux, diag/source/upper/lower out."""
    return F * 3 * S_VAL + 2 * F * S_IDX + k_in * N * S_VAL + F * S_VAL + (2 * N + 2 * F) * S_VAL


def _spmv(N, F):
    return (3 * N + 2 * F) * S_VAL + 2 * F * S_IDX


def bicgstab_iteration_bytes(N, F):
    return 2 * _spmv(N, F) + 2 * 3 * N * S_VAL + 9 * N * S_VAL + 14 * N * S_VAL


def pcg_iteration_bytes(N, F):
    return _spmv(N, F) + 3 * N * S_VAL + 5 * N * S_VAL + 9 * N * S_VAL


def _csr_spmv(n, nnz):
    """A coarse-level CSR product: values and columns, row pointers, x and y."""
    return nnz * (S_VAL + S_IDX) + (n + 1) * S_IDX + 2 * n * S_VAL


def amg_vcycle_bytes(N, F, levels):
    """One V-cycle: two Jacobi sweeps down and two up per level (each a
    product), the residual (a product and b), restriction and prolongation
    (2 n values, n aggregate indices), and the dense coarsest solve. levels
    is [(rows, CSR entries)] from level 0 (whose product is the LDU one)."""
    b = 0
    for lev, (n, nnz) in enumerate(levels[:-1]):
        sp = _spmv(N, F) if lev == 0 else _csr_spmv(n, nnz)
        b += 4 * sp + (sp + n * S_VAL) + 2 * n * S_VAL + n * S_IDX
    nc = levels[-1][0]
    return b + nc * nc * S_VAL + 2 * nc * S_VAL


def amg_pcg_iteration_bytes(N, F, levels):
    """One AMG-preconditioned CG iteration: product, dots and updates, one
    V-cycle (a solve runs one V-cycle per iteration, the first before it)."""
    return _spmv(N, F) + 5 * N * S_VAL + 9 * N * S_VAL + amg_vcycle_bytes(N, F, levels)


def amg_setup_bytes(N, F, levels):
    """Per p' solve: the Galerkin refresh of every coarse level (fine values
    and gather indices in, coarse values and inverse diagonal out) and the
    dense inverse of the coarsest level."""
    b = 0
    for lev in range(1, len(levels)):
        n, nnz = levels[lev]
        fine_vals = (N + 2 * F) if lev == 1 else levels[lev - 1][1]
        b += fine_vals * (S_VAL + S_IDX) + nnz * S_VAL + n * S_VAL
    nc = levels[-1][0]
    return b + 2 * nc * nc * S_VAL


def outer_iteration_bytes(N, F, B, turbulence):
    """Compulsory traffic of one SIMPLE outer iteration as the engine runs it
    (Krylov iterations excluded): each distinct array a stage reads, counted
    once, plus each array it writes."""
    b = 0
    b += 4 * N * S_VAL                                   # U -> U_prev copy
    b += F * 3 * S_VAL + 2 * F * S_IDX + 4 * N * S_VAL + F * S_VAL   # interpolated flux
    n_grad = 5 if turbulence else 3                      # U(2), p (+k, omega)
    b += n_grad * _gg(N, F, B)
    if turbulence:
        b += 17 * N * S_VAL                              # SST closure sweep
    b += 2 * _assembly(N, F, 9)                          # x/y momentum
    b += _gg(N, F, B)                                    # grad Ux for the y transpose stress
    b += _assembly(N, F, 7)                              # Rhie-Chow flux + p' assembly
    b += _gg(N, F, B) + 9 * N * S_VAL                    # grad p', U/p correction
    b += F * 3 * S_VAL + 2 * F * S_IDX + 2 * N * S_VAL   # flux correction
    if turbulence:
        b += _assembly(N, F, 8) + _assembly(N, F, 10)    # k, omega
        b += 2 * 3 * N * S_VAL                           # clamps and change norms
    b += 4 * N * S_VAL                                   # convergence norms
    return b


def solve_bytes(N, F, B, iters, krylov, turb_start, settings, levels=None):
    """Algorithmic bytes of one solve from its executed iteration counts
    (levels: the AMG hierarchy when the pressure solver is AMG)."""
    turb_iters = 0
    if iters > turb_start:
        turb_iters = (iters - turb_start + settings.turb_update_interval - 1) \
            // settings.turb_update_interval
    plain_iters = iters - turb_iters
    b = plain_iters * outer_iteration_bytes(N, F, B, False)
    b += turb_iters * outer_iteration_bytes(N, F, B, True)
    per = {"BiCGSTAB": bicgstab_iteration_bytes(N, F), "PCG": pcg_iteration_bytes(N, F)}
    if levels:
        per["AMG"] = amg_pcg_iteration_bytes(N, F, levels)
        b += iters * amg_setup_bytes(N, F, levels)  # one p' solve per outer iteration
    solvers = [settings.momentum_solver, settings.momentum_solver, settings.pressure_solver,
               settings.turbulence_solver, settings.turbulence_solver]
    for count, name in zip(krylov, solvers):
        b += int(count) * per[name]
    return b


# --------------------------------------------------------------- helpers
def _host_cpus():
    """CPUs this process may run on (its affinity mask)."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def _cores():
    """Workers for the CPU baselines: every CPU of the process's affinity
    mask, within the CPU share the host grants the job (OMP_NUM_THREADS when
    it is set: on a shared GPU host more workers than the share would only
    queue behind each other)."""
    n = _host_cpus()
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def _host():
    """What the host offers the CPU baselines: its CPU count, this process's
    affinity mask and the job's CPU share (OMP_NUM_THREADS)."""
    return {"nproc": os.cpu_count(), "affinity_cpus": _host_cpus(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _traffic_from_profile(name="simple_solve_kernel_channel_sweep64_traffic.json"):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a copy of known size)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def cpu_baseline(prob, theta, cores):
    """The CPU FP64 oracle, one draw per thread (the ctypes call releases the
    GIL), on `cores` draws of the sweep."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    O = oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs, prob.nu,
                                prob.init_vector())
    oracle_py.lib()
    rows = theta[:cores]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        res = list(ex.map(lambda th: O.evaluate(prob.coefficients(th)), rows))
    dt = time.perf_counter() - t0
    conv = sum(1 for r in res if r["status"] == 0)
    return res, {"value": conv / dt, "unit": "converged solves/s", "cores": cores,
                 "kind": "port",
                 "sample": f"{len(rows)} draws of the 64-draw sweep, one per thread, "
                           f"{dt:.1f} s wall"}


def reference_cpu_baseline(prob, theta, cores):
    """The reference's own solver, compiled from its sources into
    oracle/_ref/ref_harness (where that was built), one draw per process on
    `cores` processes at once. Returns (rows, baseline) like cpu_baseline, or
    None when the harness is absent."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.access(harness, os.X_OK):
        return None
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_fixtures import problem_lines
    tmp = tempfile.mkdtemp(prefix="rq_ref_")
    mesh_path = os.path.join(tmp, "mesh.msh")
    prob.mesh.save_binary(mesh_path)
    rows = theta[:cores]
    paths = []
    for i, th in enumerate(rows):
        path = os.path.join(tmp, f"case{i}.txt")
        with open(path, "w") as fh:
            fh.write("\n".join(problem_lines(prob, mesh_path, [th])) + "\n")
        paths.append(path)
    t0 = time.perf_counter()
    procs = [subprocess.Popen([harness, p], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                              text=True) for p in paths]
    outs = [pr.communicate()[0] for pr in procs]
    dt = time.perf_counter() - t0
    if any(pr.returncode != 0 for pr in procs):
        return None
    res = []
    for out in outs:
        body = out[out.index("{"):].replace("NaN", "null").replace("-Infinity", "-1e999")
        row = json.loads(body.replace("Infinity", "1e999"))["rows"][0]
        res.append({"status": int(row["status"]), "predictions":
                    np.asarray([np.nan if v is None else v for v in row["predictions"]], float),
                    "wall_seconds": row.get("wall_seconds")})
    conv = sum(1 for r in res if r["status"] == 0)
    solve_s = sum(r["wall_seconds"] or 0.0 for r in res)
    return res, {"value": conv / dt, "unit": "converged solves/s", "cores": len(rows),
                 "kind": "reference",
                 "sample": f"{len(rows)} draws of the 64-draw sweep, one per process of the "
                           f"reference's solver (oracle/_ref/ref_harness), {dt:.1f} s wall "
                           f"({solve_s / max(len(rows), 1):.2f} s per solve inside the solver)"}


def single_solve_leg(cpu):
    """The single FP64 solve at the default coefficients (a1 0.31, beta* 0.09)
    on the 1x128 channel: one launch on the GPU (the plain method), and the
    CPU oracle on one thread beside it."""
    from cfdr_amd import cases, engine
    prob = cases.channel_problem()
    theta = np.asarray([[0.31, 0.09]])
    b = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                            prob.u_init, prob.p_init, prob.k_init, prob.omega_init)
    b.evaluate(theta)  # first launch loads the code object
    t0 = time.perf_counter()
    r = b.evaluate(theta)
    dt = time.perf_counter() - t0
    out = {"status": r.statuses[0].name, "iterations": int(r.iteration_counts[0]),
           "gpu_ms": round(1e3 * dt, 1), "kernel_ms": round(r.kernel_ms, 1)}
    if cpu:
        ref = reference_cpu_baseline(prob, theta, 1)
        if ref is not None:
            o = ref[0][0]
            out["cpu_baseline"] = {"value_ms": round(1e3 * o["wall_seconds"], 1), "cores": 1,
                                   "kind": "reference",
                                   "sample": "the same solve by the reference's solver "
                                             "(oracle/_ref/ref_harness), its own solve timer"}
        else:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle_py
            O = oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs, prob.nu,
                                        prob.init_vector())
            t0 = time.perf_counter()
            o = O.evaluate(prob.coefficients(theta[0]))
            out["cpu_baseline"] = {"value_ms": round(1e3 * (time.perf_counter() - t0), 1),
                                   "cores": 1, "kind": "port",
                                   "sample": "the same solve on one thread"}
        if o["status"] != int(r.statuses[0]):
            raise SystemExit("bench.py: the single solve's status differs from the CPU's")
        out["max_rel_err_vs_cpu_baseline"] = float(
            np.max(np.abs(r.predictions[0] - o["predictions"]) / np.abs(o["predictions"])))
    b.close()
    print(f"# single solve: {out['gpu_ms']} ms on the GPU", file=sys.stderr, flush=True)
    return out


def copy_calibration(mib):
    """Three device-to-device copies of `mib` MiB through the HIP runtime."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), n) != 0:
            raise SystemExit("bench.py: hipMalloc failed")
    if hip.hipMemset(a, 1, n) != 0:
        raise SystemExit("bench.py: hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, n, 3) != 0:  # hipMemcpyDeviceToDevice
            raise SystemExit("bench.py: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": mib << 20, "copies": 3}), flush=True)
    return 0


HILL_MESHES = ((60, 32), (120, 64), (240, 128))


def served_leg(n_requests, rate, hill_cap=0, plain_check=4):
    """The served stream: seeded Poisson arrivals of single-solve requests on
    1D channels of 64/128/256/512 cells (one first-cell height), submitted at
    their arrival times to the device request loop; latency = completion -
    arrival. The plain path serves the same arrivals one launch at a time, in
    order, with each mesh's solve time measured alone on the same GPU. With
    hill_cap > 0 the meshes also include the 60x32, 120x64 and 240x128 hills,
    whose solves are capped at hill_cap outer iterations. The plain path's
    latencies come from that model (queueing in arrival order, each mesh's
    solve time measured once); its first plain_check requests are also
    served for real, one launch each at their arrival times, and reported
    beside the model's latencies for the same requests."""
    from cfdr_amd import cases, engine
    from cfdr_amd.serve import Server
    keys = list(cases.SERVED_CHANNELS)
    probs = {n: cases.channel_problem(ny=n, ratio=cases.channel_ratio(n))
             for n in cases.SERVED_CHANNELS}
    if hill_cap > 0:
        for nx, ny in HILL_MESHES:
            key = f"hill{nx}x{ny}"
            keys.append(key)
            probs[key] = cases.hill_problem(nx=nx, ny=ny, max_iterations=hill_cap)
    stream = cases.served_stream(n_requests, rate, meshes=tuple(keys))

    def row(key, th):
        # hill problems infer (a1, beta*, sigma_k1, sigma_w1): the stream's
        # (a1, beta*) with sigma_k1, sigma_w1 at the centre of the hill draws
        return np.asarray([th[0], th[1], 0.85, 0.5]) if isinstance(key, str) else th
    done = {}
    with Server(max_slots=256, idle_exit_seconds=5.0) as srv:
        pid = {n: srv.add_problem(p.mesh, p.params, p.obs, p.bcs, p.nu, p.settings, p.u_init,
                                  p.p_init, p.k_init, p.omega_init) for n, p in probs.items()}
        t0 = time.perf_counter()
        tickets = []

        def collect():
            now = time.perf_counter() - t0
            for k, (ta, t, n) in enumerate(tickets):
                if k not in done:
                    r = srv.poll(t)
                    if r is not None:
                        done[k] = (ta, now, r, n)

        for ta, n, th in stream:
            while time.perf_counter() - t0 < ta:
                collect()
                time.sleep(2e-4)
            tickets.append((ta, srv.submit(pid[n], row(n, th)), n))
        print(f"# served: {len(tickets)} requests submitted", file=sys.stderr, flush=True)
        deadline = time.perf_counter() + 600
        while len(done) < len(tickets) and time.perf_counter() < deadline:
            collect()
            time.sleep(5e-4)
        stats = srv.stats()
    if len(done) < len(tickets):
        raise SystemExit("bench.py: served requests did not finish within 600 s")
    lat = np.array([done[k][1] - done[k][0] for k in range(len(tickets))])
    conv = sum(1 for k in done if done[k][2].status == engine.EvaluationStatus.Converged)
    span = max(d[1] for d in done.values()) - min(d[0] for d in done.values())
    # plain: one launch per request, in arrival order, solve times measured alone
    svc = {}
    for n, p in probs.items():
        b = engine.ForwardBatch(p.mesh, p.params, p.obs, p.bcs, p.nu, p.settings, p.u_init,
                                p.p_init, p.k_init, p.omega_init)
        first = next(th for ta, m, th in stream if m == n)
        t1 = time.perf_counter()
        b.evaluate(np.asarray([row(n, first)]))
        svc[n] = time.perf_counter() - t1
        b.close()
    end = 0.0
    plat = []
    for ta, n, th in stream:
        end = max(end, ta) + svc[n]
        plat.append(end - ta)
    plat = np.array(plat)
    pspan = end - stream[0][0]
    check = None
    if plain_check > 0:
        # the first requests served for real on the plain path: engines made
        # before the clock starts, one launch per request in arrival order
        sub = stream[:plain_check]
        engines = {n: engine.ForwardBatch(probs[n].mesh, probs[n].params, probs[n].obs,
                                          probs[n].bcs, probs[n].nu, probs[n].settings,
                                          probs[n].u_init, probs[n].p_init, probs[n].k_init,
                                          probs[n].omega_init) for n in {m for _, m, _ in sub}}
        measured = []
        try:
            t0 = time.perf_counter()
            for ta, n, th in sub:
                while time.perf_counter() - t0 < ta:
                    time.sleep(1e-3)
                engines[n].evaluate(np.asarray([row(n, th)]))
                measured.append(time.perf_counter() - t0 - ta)
        finally:
            for b in engines.values():
                b.close()
        model = plat[:plain_check]
        check = {"requests": [str(n) for _, n, _ in sub],
                 "measured_ms": [round(1e3 * v, 1) for v in measured],
                 "model_ms": [round(1e3 * float(v), 1) for v in model],
                 "max_rel_diff": round(float(np.max(np.abs(np.asarray(measured) - model) /
                                                    model)), 4)}
        print(f"# served: plain check on {plain_check} requests, max rel diff "
              f"{check['max_rel_diff']}", file=sys.stderr, flush=True)
    by_status = {}
    for k in done:
        name = done[k][2].status.name
        by_status[name] = by_status.get(name, 0) + 1
    return {"value": round(len(stream) / span, 4), "unit": "completed solves/s",
            "requests": len(stream), "rate_per_s": rate, "statuses": by_status,
            "converged_per_s": round(conv / span, 4),
            "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 1),
                           "p99": round(1e3 * float(np.percentile(lat, 99)), 1)},
            "slots": stats["slots"],
            "meshes": "1D channels of 64/128/256/512 cells" +
                      (f" and 60x32/120x64/240x128 hills (capped at {hill_cap} outer "
                       "iterations)" if hill_cap > 0 else "") +
                      ", uniform; theta uniform a1 in [0.28, 0.34], beta* in [0.084, 0.096] "
                      "(hills: sigma_k1, sigma_w1 at the hill defaults); seed 4242",
            "plain": {"value": round(len(stream) / pspan, 4), "unit": "completed solves/s",
                      "latency_ms": {"p50": round(1e3 * float(np.percentile(plat, 50)), 1),
                                     "p99": round(1e3 * float(np.percentile(plat, 99)), 1)},
                      "solve_s_by_cells": {str(n): round(v, 3) for n, v in svc.items()},
                      "note": "same arrivals served in order, one launch each; each mesh's "
                              "solve time measured alone",
                      "measured_check": check},
            "speedup_vs_plain": round((len(stream) / span) / (len(stream) / pspan), 2),
            "note": "tol 1e-8 and the 150000-iteration cap of the single-solve channel: the "
                    "256- and 512-cell channels end Unconverged at the cap on every path"}


def _mixed_vs_fp64(rm, r):
    """FP32-Krylov rows against the FP64 rows of the same batch: status
    agreement and the largest relative probe difference of rows that both
    report predictions for, over all of them and over the rows both runs
    converge (the rows a posterior admits; a row stopped at the iteration
    cap holds an unfinished iterate, which depends on the inner solves'
    iterates and not only on the fixed point)."""
    def err(mask):
        return (float(np.max(np.abs(rm.predictions[mask] - r.predictions[mask]) /
                             np.abs(r.predictions[mask]))) if mask.any() else None)
    ok = np.isin(r.status_codes, (0, 1)) & (rm.status_codes == r.status_codes)
    conv = (r.status_codes == 0) & (rm.status_codes == 0)
    return {"statuses_match_fp64": bool(np.array_equal(rm.status_codes, r.status_codes)),
            "converged_flags_match_fp64": bool(np.array_equal(rm.convergence_mask,
                                                              r.convergence_mask)),
            "max_rel_probe_diff_vs_fp64": err(ok),
            "max_rel_probe_diff_vs_fp64_converged_rows": err(conv)}


def hills_leg(full, cores, cpu, meshes=("hill_120x64", "hill_480x256"), block_size=0,
              grid_blocks=0, plain_launch=False):
    """Hill meshes: the 16-draw 120x64 batch (AMG pressure, tol 1e-4) and the
    8-draw 480x256 batch, each timed per outer iteration on solves capped at
    a few hundred (resp. ten) iterations, on the launch shape the autotuner
    picks (timed on a shorter capped run first), in FP64 and with FP32 Krylov
    and FP64 defect correction. The CPU oracle runs the 120x64 draws one per
    thread with the same cap. With full=True the 16-draw 120x64 batch also
    runs to convergence in both precisions (cap 25000; minutes). meshes picks
    the capped batches; block_size and grid_blocks (both nonzero) pin the
    launch shape instead of tuning it, and plain_launch dispatches the
    multi-workgroup grid as an ordinary launch, for profiler runs."""
    from cfdr_amd import cases, engine
    out = {}
    for key, nx, ny, nd, cap, tune_its in (("hill_120x64", 120, 64, 16, 200, 20),
                                           ("hill_480x256", 480, 256, 8, 10, 2)):
        if key not in meshes:
            continue
        prob = cases.hill_problem(nx=nx, ny=ny, max_iterations=cap)
        th = cases.hill_draws(nd)
        legs = {}
        for prec in ("fp64", "mixed"):
            pinned = block_size > 0 and grid_blocks > 0
            b = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu,
                                    prob.settings, prob.u_init, prob.p_init, prob.k_init,
                                    prob.omega_init, mixed_precision=prec == "mixed",
                                    block_size=block_size if pinned else 0,
                                    grid_blocks=grid_blocks if pinned else 0,
                                    grid_plain_launch=plain_launch)
            tune = ({"block_size": block_size, "grid_blocks": grid_blocks, "candidates": []}
                    if pinned else b.autotune(th, iterations=tune_its))
            legs[prec] = (b.evaluate(th), b.last_timing(), tune)
            legs["levels"] = b.amg_levels()
            b.close()
            print(f"# {key} {prec}: {legs[prec][0].kernel_ms / cap:.2f} ms per outer iteration "
                  f"(tuned block {tune['block_size']}, {tune['grid_blocks']} per draw)",
                  file=sys.stderr, flush=True)
        r, t, tune = legs["fp64"]
        rm, tm, tunem = legs["mixed"]
        kry = np.asarray(r.krylov_iterations, dtype=float).mean(axis=0) / cap
        m = prob.mesh
        launch_bytes = sum(solve_bytes(m.n_cells, m.n_internal, m.n_boundary,
                                       int(r.iteration_counts[i]), r.krylov_iterations[i],
                                       prob.settings.turb_start_iter, prob.settings,
                                       legs["levels"])
                           for i in range(nd))
        achieved = launch_bytes / (r.kernel_ms / 1e3) / 1e9
        out[key] = {"draws": nd, "cells": prob.mesh.n_cells, "capped_iterations": cap,
                    "ms_per_outer_iteration": round(r.kernel_ms / cap, 3),
                    "draw_iterations_per_s": round(nd * cap / (r.kernel_ms / 1e3), 1),
                    "grid_blocks": t["grid_blocks"], "block_size": t["block_size"],
                    "autotune": {"iterations": tune_its, "candidates": tune["candidates"]},
                    "krylov_per_outer_iteration": [round(float(k), 1) for k in kry],
                    "roofline": {"bound": "hbm", "achieved": round(achieved, 2),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(achieved / HBM_PEAK_GBS, 5),
                                 "kernel": "grid_solve_kernel" if t["grid_blocks"] > 1
                                 else "simple_solve_kernel",
                                 "algorithmic_bytes_per_launch": int(launch_bytes),
                                 "traffic": (_traffic_from_profile(
                                     "grid_solve_kernel_hill120x64_traffic.json")
                                     if key == "hill_120x64" else None),
                                 "traffic_note": ("HBM bytes of one such launch measured at 16 "
                                                  "workgroups per draw (profiles/)"
                                                  if key == "hill_120x64" else None),
                                 "amg_levels": legs["levels"]},
                    "mixed": dict({"ms_per_outer_iteration": round(rm.kernel_ms / cap, 3),
                                   "grid_blocks": tm["grid_blocks"],
                                   "block_size": tm["block_size"],
                                   "at": f"the {cap}-iteration cap"},
                                  **_mixed_vs_fp64(rm, r))}
        if cpu and key == "hill_120x64":
            ccap = cap  # the same capped window as the GPU rate
            prob.settings.max_iterations = ccap
            rows = np.asarray([th[i % nd] for i in range(cores)])
            ref = reference_cpu_baseline(prob, rows, cores)
            if ref is not None:
                # the reference's solver, one draw per process; its own solve timers
                per = [r["wall_seconds"] for r in ref[0]]
                dt = max(per)
                out[key]["cpu_baseline"] = {
                    "value": round(cores * ccap / dt, 1), "unit": "draw-iterations/s",
                    "cores": cores, "kind": "reference", "host": _host(),
                    "sample": f"{cores} draws capped at {ccap} outer iterations, one per "
                              f"process of the reference's solver, slowest {dt:.1f} s"}
                one = reference_cpu_baseline(prob, rows[:1], 1)
                if one is not None:
                    t1 = one[0][0]["wall_seconds"]
                    out[key]["cpu_baseline_1worker"] = {
                        "value": round(ccap / t1, 1), "unit": "draw-iterations/s", "cores": 1,
                        "kind": "reference",
                        "sample": f"1 draw capped at {ccap} outer iterations, one process of "
                                  f"the reference's solver, {t1:.1f} s"}
            else:
                sys.path.insert(0, os.path.join(ROOT, "oracle"))
                import oracle_py
                O = oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs,
                                            prob.nu, prob.init_vector())
                t0 = time.perf_counter()
                with ThreadPoolExecutor(max_workers=cores) as ex:
                    list(ex.map(lambda v: O.evaluate(prob.coefficients(v)), rows))
                dt = time.perf_counter() - t0
                out[key]["cpu_baseline"] = {
                    "value": round(cores * ccap / dt, 1), "unit": "draw-iterations/s",
                    "cores": cores, "kind": "port", "host": _host(),
                    "sample": f"{cores} draws capped at {ccap} outer iterations, one per "
                              f"thread, {dt:.1f} s wall"}
    if full:
        prob = cases.hill_problem(nx=120, ny=64)
        th = cases.hill_draws(16)
        full_rows = {}
        for prec in ("fp64", "mixed"):
            b = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu,
                                    prob.settings, prob.u_init, prob.p_init, prob.k_init,
                                    prob.omega_init, mixed_precision=prec == "mixed")
            print(f"# hill_120x64 {prec} to convergence (cap 25000) ...", file=sys.stderr,
                  flush=True)
            t0 = time.perf_counter()
            r = b.evaluate(th)
            dt = time.perf_counter() - t0
            b.close()
            full_rows[prec] = r
            out["hill_120x64_full" + ("" if prec == "fp64" else "_mixed")] = {
                "draws": 16, "seconds": round(dt, 1),
                "converged_per_s": round(int(np.sum(r.convergence_mask)) / dt, 4),
                "completed_per_s": round(16 / dt, 4),
                "statuses": {s.name: int(sum(1 for x in r.statuses if x == s))
                             for s in set(r.statuses)},
                "iterations": [int(i) for i in r.iteration_counts]}
        out["hill_120x64_full_mixed"].update(_mixed_vs_fp64(full_rows["mixed"],
                                                            full_rows["fp64"]))
    return out


def warm_leg(steps):
    """Warm starts on the 1x32 channel (ratio 1.15, tol 1e-6): a converged base
    point (0.31, 0.09), then a pattern-search poll of its 4 neighbours at a
    relative step of 1 % (and of 10 %), solved cold in one batch launch and
    warm-started from the base state in one launch (rq_model_evaluate_batch).
    Each poll is repeated `steps` times; every warm poll starts from a model
    holding only the base state. (On the 1x128 channel at tol 1e-8 the
    reference classifies such warm starts as diverged, and so does this
    engine.)"""
    from cfdr_amd import cases, engine
    from cfdr_amd.model import ForwardModel
    prob = cases.channel_problem(ny=32, ratio=1.15, n_probes=30, tol=1e-6,
                                 max_iterations=40000)
    base = np.asarray([0.31, 0.09])
    args = (prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings, prob.u_init,
            prob.p_init, prob.k_init, prob.omega_init)
    batch = engine.ForwardBatch(*args)
    out = {}
    for label, rel in (("poll_1pct", 0.01), ("poll_10pct", 0.10)):
        stencil = np.asarray([base * [1 + rel, 1], base * [1 - rel, 1], base * [1, 1 + rel],
                              base * [1, 1 - rel]])
        cold_s, warm_s = [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            rc = batch.evaluate(stencil)
            cold_s.append(time.perf_counter() - t0)
            m = ForwardModel(*args)
            m.evaluate(base)
            t0 = time.perf_counter()
            rw, used = m.evaluate_batch(stencil)
            warm_s.append(time.perf_counter() - t0)
            m.close()
        ok = (rc.status_codes == 0) & (rw.status_codes == 0)
        err = float(np.max(np.abs(rw.predictions[ok] - rc.predictions[ok]) /
                           np.abs(rc.predictions[ok]))) if ok.any() else None
        out[label] = {
            "rows": len(stencil), "warm_started": int(np.sum(used)),
            "cold_ms": round(1e3 * statistics.mean(cold_s), 1),
            "warm_ms": round(1e3 * statistics.mean(warm_s), 1),
            "speedup": round(statistics.mean(cold_s) / statistics.mean(warm_s), 2),
            "converged_cold": int(np.sum(rc.convergence_mask)),
            "converged_warm": int(np.sum(rw.convergence_mask)),
            "cold_iterations_mean": float(np.mean(rc.iteration_counts)),
            "warm_iterations_mean": float(np.mean(rw.iteration_counts)),
            "statuses_cold": [s.name for s in rc.statuses],
            "statuses_warm": [s.name for s in rw.statuses],
            "max_rel_diff_warm_vs_cold": err}
        print(f"# warm {label}: cold {out[label]['cold_ms']} ms, warm {out[label]['warm_ms']} ms",
              file=sys.stderr, flush=True)
    batch.close()
    return out


def warm_fd_leg(cpu):
    """The warm-FD Jacobian (eleven SST coefficients, 22 perturbed solves
    started from the converged base) on the 1x128 channel at tol 1e-8: the
    GPU runs the 22 solves in one launch; the CPU oracle runs them one after
    another, as the reference does."""
    from cfdr_amd import cases
    from cfdr_amd.config import COEFF_DEFAULTS
    from cfdr_amd.sensitivity import ParameterSensitivity
    prob = cases.channel_problem()
    theta = np.asarray(COEFF_DEFAULTS)
    s = ParameterSensitivity(prob.mesh, prob.obs, prob.bcs, prob.nu, prob.settings, prob.u_init,
                             prob.p_init, prob.k_init, prob.omega_init)
    if s.solve_state(theta) != 0:
        raise SystemExit("bench.py: the warm-FD base state did not converge")
    t0 = time.perf_counter()
    g = s.eta_jacobian_warm_fd(theta)
    gpu_s = time.perf_counter() - t0
    kernel_ms = s.last_kernel_ms()
    s.close()
    out = {"solves": 22, "columns_converged": int(np.sum(g.krylov_converged)),
           "simple_iterations": int(g.n_residual_evals), "gpu_ms": round(1e3 * gpu_s, 1),
           "kernel_ms": round(kernel_ms, 1)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        O = oracle_py.OracleProblem(prob.mesh, prob.bcs, prob.settings, prob.obs, prob.nu,
                                    prob.init_vector())
        st, state = oracle_py.solve_state(O, theta)
        t0 = time.perf_counter()
        r = oracle_py.eta_jacobian_warm_fd(O, state, theta)
        seq = time.perf_counter() - t0
        scale = np.maximum(np.max(np.abs(r["d_obs_d_theta"]), axis=0), 1e-3)
        out["cpu_baseline"] = {"value_ms": round(1e3 * seq, 1), "cores": 1, "kind": "port",
                               "sample": "the 22 warm solves one after another (the reference's "
                                         "order)"}
        out["max_scaled_diff_vs_oracle"] = float(np.max(np.abs(g.d_obs_d_theta -
                                                               r["d_obs_d_theta"]) / scale))
        out["speedup_vs_cpu_sequential"] = round(seq / gpu_s, 2)
    print(f"# warm-FD Jacobian: {out['gpu_ms']} ms on the GPU", file=sys.stderr, flush=True)
    return out


def calibration_leg():
    """The staged calibration (uniform design, pattern search, 5x5 Laplace
    grid, two 64-draw importance posteriors, split conformal) on the 1x32
    geometric channel (ratio 1.15, tol 1e-6), each stage one batched launch.
    The observed profile is the engine's own prediction at theta* = (0.322,
    0.0952) at 30 wall distances (y+ 1 to 182), noise 3e-4 of each value,
    every third point in the likelihood; the stage-1 box is centred on
    theta* (a1 +/- 0.004, beta* +/- 0.0012). On the 1x128 channel the same
    synthetic likelihood is a ridge narrower than the final pattern-search
    step, so its 5x5 grid is not quadratic and the Laplace stage refuses it
    (the peak it fits lies outside the grid), as the reference's driver does
    (calibration_coverage.py:253-254). The reference's study runs 241 solves
    one per thread."""
    from cfdr_amd import cases, engine
    from cfdr_amd.calibration import Calibration, CalibrationSettings, channel_batch
    prob = cases.channel_problem(ny=32, ratio=1.15, n_probes=30, tol=1e-6, max_iterations=40000)
    y = np.array([ob.location[1] for ob in prob.obs.observables])
    gen = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                              prob.u_init, prob.p_init, prob.k_init, prob.omega_init)
    theta_star = np.array([0.322, 0.0952])
    truth = gen.evaluate(theta_star[None, :]).predictions[0]
    gen.close()
    group = np.arange(truth.size) % 3
    sigma = 3e-4 * truth
    batch = channel_batch(prob, y, truth, np.where(group == 0, sigma, np.inf))
    settings = CalibrationSettings(stage1_box=((0.318, 0.094), (0.326, 0.0964)))
    t0 = time.perf_counter()
    try:
        res = Calibration.for_batch(batch, settings).run(truth, sigma, group)
    finally:
        batch.close()
    dt = time.perf_counter() - t0
    print(f"# calibration: {res.solves} solves in {res.launches} launches, {dt:.1f} s",
          file=sys.stderr, flush=True)
    return {"solves": res.solves, "launches": res.launches, "seconds": round(dt, 2),
            "solves_per_s": round(res.solves / dt, 3),
            "optimum": [float(v) for v in res.optimum],
            "laplace_center": [float(v) for v in res.center],
            "posterior_mean": [float(v) for v in res.main.theta_mean],
            "effective_sample_size": round(float(res.main.effective_sample_size), 2),
            "held_out_covered": {k: res.coverage[k] for k in
                                 ("held_out", "standard_covered", "recalibrated_covered",
                                  "replicate_covered")},
            "workload": "1x32 channel (ratio 1.15, tol 1e-6), synthetic profile at theta* "
                        "(0.322, 0.0952), 30 points, sigma 3e-4 relative, stage-1 box theta* "
                        "+/- (0.004, 0.0012)",
            "reference": "docs/figures/calibration_coverage.py runs its 241 solves in about 50 "
                         "minutes on four threads (its 3x128 periodic channel)"}


PHASE_NAMES = ("gradients + SST closure", "x-momentum rows", "x solve", "y-momentum rows",
               "y solve", "Rhie-Chow and p' rows", "p' solve", "corrections", "k rows",
               "k solve", "omega rows", "omega solve", "norms", "bookkeeping")


def phase_clocks(block_size, max_iterations, mixed):
    """Per-phase shader clocks of draw 0 of the sweep (rq_settings.profile):
    one draw, capped at max_iterations outer iterations, one JSON line."""
    from cfdr_amd import cases, engine
    prob = cases.channel_problem(max_iterations=max_iterations)
    theta = cases.channel_draws(64)[:1]
    b = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu, prob.settings,
                            prob.u_init, prob.p_init, prob.k_init, prob.omega_init,
                            block_size=block_size, profile=True, mixed_precision=mixed)
    try:
        r = b.evaluate(theta)
    finally:
        b.close()
    it = int(r.iteration_counts[0])
    cyc = np.asarray(r.phase_cycles[0][:len(PHASE_NAMES)], dtype=float)
    print(json.dumps({"workload": "draw 0 of the 64-draw channel sweep, 1x128 cells",
                      "block_size": block_size, "mixed_precision": bool(mixed),
                      "iterations": it, "kernel_ms": round(float(r.kernel_ms), 2),
                      "us_per_iteration": round(float(r.kernel_ms) * 1e3 / it, 2),
                      "clocks_per_iteration": round(float(cyc.sum()) / it),
                      "krylov_iterations": [int(v) for v in r.krylov_iterations[0]],
                      "phases": {n: round(float(c) / it) for n, c in zip(PHASE_NAMES, cyc)}}))
    return 0


def _leg(name, fn, *args):
    """Run one auxiliary leg. An exception raised by the leg's own logic (a
    calibration stage that refuses its grid, a CPU baseline that cannot
    start) is reported in that leg's entry instead of ending the run, so the
    headline line is always printed."""
    try:
        return fn(*args)
    except Exception as exc:  # noqa: BLE001 - reported in the line
        print(f"# {name} leg failed: {type(exc).__name__}: {exc}", file=sys.stderr, flush=True)
        return {"error": f"{type(exc).__name__}: {exc}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--draws", type=int, default=64)
    ap.add_argument("--plain-sample", type=int, default=4)
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the engine's default launch shape for the sweep")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the sweep directly instead of replaying a captured HIP graph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=16.0)
    ap.add_argument("--serve-hills", type=int, default=0, metavar="CAP",
                    help="add the 60x32/120x64/240x128 hills to the served stream, their solves "
                         "capped at CAP outer iterations")
    ap.add_argument("--serve-plain-check", type=int, default=4, metavar="K",
                    help="also serve the stream's first K requests on the plain path for real "
                         "and report them beside the plain path's model (0: skip)")
    ap.add_argument("--max-iterations", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--no-hills", action="store_true")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single default-coefficient solve (GPU and one CPU thread)")
    ap.add_argument("--no-warm", action="store_true")
    ap.add_argument("--no-warm-fd", action="store_true")
    ap.add_argument("--no-hills-full", action="store_true",
                    help="skip the 16-draw 120x64 hill batch solved to convergence (FP64 and "
                         "mixed, about 3.5 minutes together)")
    ap.add_argument("--hill-meshes", default="hill_120x64,hill_480x256",
                    help="capped hill batches to time (comma list)")
    ap.add_argument("--hills-block-size", type=int, default=0,
                    help="with --hills-grid-blocks: pin the hill launch shape (profiler runs)")
    ap.add_argument("--hills-grid-blocks", type=int, default=0)
    ap.add_argument("--hills-plain-launch", action="store_true",
                    help="launch the multi-workgroup hill grids as ordinary dispatches (the "
                         "profiler writes no trace of a process that made cooperative launches)")
    ap.add_argument("--no-calibration", action="store_true",
                    help="skip the staged calibration on the 1x128 channel")
    ap.add_argument("--phase-clocks", action="store_true",
                    help="print the per-phase shader clocks of one draw (capped at "
                         "--max-iterations, default 3000) at --block-size and exit")
    ap.add_argument("--mixed", action="store_true",
                    help="with --phase-clocks: FP32 Krylov with FP64 defect correction")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.phase_clocks:
        return phase_clocks(args.block_size or 256, args.max_iterations or 3000, args.mixed)
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(draws are independent: several GPUs would each run their own batch)",
              file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0:
        print("bench.py: --steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        return 2

    from cfdr_amd import cases, engine

    prob = cases.channel_problem()
    if args.max_iterations:
        prob.settings.max_iterations = args.max_iterations
    theta = cases.channel_draws(args.draws)
    mesh = prob.mesh
    N, F, B = mesh.n_cells, mesh.n_internal, mesh.n_boundary
    batch = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu,
                                prob.settings, prob.u_init, prob.p_init, prob.k_init,
                                prob.omega_init, block_size=args.block_size)
    print(f"# workload: {args.draws} draws, {N} cells, build {engine.build_info()}",
          file=sys.stderr, flush=True)
    tune = None
    if not args.no_autotune and not args.block_size:
        # launch shape from the autotuner, timed on the sweep capped at 2000
        # outer iterations (untimed: tuning happens once per problem)
        tune = batch.autotune(theta, iterations=2000)
        print(f"# autotune: block {tune['block_size']}, {tune['grid_blocks']} workgroup(s) per "
              f"draw", file=sys.stderr, flush=True)

    if not args.no_graph:
        # the sweep's launch is captured on its first evaluate (a warmup step)
        # and replayed by every later one
        batch.set_graph_replay(True)
    for w in range(args.warmup):
        batch.evaluate(theta)
        print(f"# warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    # evaluate() returns after its stream has synchronised: the timed region
    # starts and ends with the device idle.
    step_ms, kernel_ms, results = [], [], []
    t_all = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        r = batch.evaluate(theta)
        step_ms.append(1e3 * (time.perf_counter() - t0))
        kernel_ms.append(r.kernel_ms)
        results.append(r)
        print(f"# step {k + 1}/{args.steps}: {step_ms[-1]:.1f} ms", file=sys.stderr, flush=True)
    total_s = time.perf_counter() - t_all
    last = results[-1]
    converged = int(sum(int(np.sum(r.convergence_mask)) for r in results))
    value = converged / total_s
    if args.max_iterations:
        print("# capped solves: value counts converged rows only", file=sys.stderr, flush=True)
    for r in results[1:]:
        if not np.array_equal(r.predictions, last.predictions, equal_nan=True):
            raise SystemExit("bench.py: repeated steps differ")

    # roofline of the solve kernel (the only kernel of a step)
    launch_bytes = sum(solve_bytes(N, F, B, int(last.iteration_counts[i]),
                                   last.krylov_iterations[i], prob.settings.turb_start_iter,
                                   prob.settings)
                       for i in range(len(theta)) if last.status_codes[i] in (0, 1))
    kernel_avg_s = statistics.mean(kernel_ms) / 1e3
    achieved = launch_bytes / kernel_avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": _traffic_from_profile(),
                "traffic_note": ("HBM bytes of one launch of this kernel at 256 threads, from "
                                 "rocprofv3 FETCH_SIZE and WRITE_SIZE scaled by a 256 MiB copy "
                                 "(profiles/); achieved counts algorithmic bytes, and the draws' "
                                 "state stays in LDS"),
                "kernel": "simple_solve_kernel", "kernel_ms": round(1e3 * kernel_avg_s, 3),
                "algorithmic_bytes_per_launch": int(launch_bytes)}

    # the same sweep with FP32 Krylov iterations and FP64 defect correction
    mixed = None
    if not args.no_mixed:
        mb = engine.ForwardBatch(prob.mesh, prob.params, prob.obs, prob.bcs, prob.nu,
                                 prob.settings, prob.u_init, prob.p_init, prob.k_init,
                                 prob.omega_init, block_size=args.block_size,
                                 mixed_precision=True)
        t0 = time.perf_counter()
        rm = mb.evaluate(theta)
        dt = time.perf_counter() - t0
        same = bool(np.array_equal(rm.status_codes, last.status_codes))
        ok = (last.status_codes == 0) & (rm.status_codes == 0)
        err_fp64 = float(np.max(np.abs(rm.predictions[ok] - last.predictions[ok]) /
                                np.abs(last.predictions[ok]))) if ok.any() else None
        mixed = {"value": round(int(np.sum(rm.convergence_mask)) / dt, 4),
                 "unit": "converged solves/s", "dtype": "f64 assembly/norms, f32 Krylov",
                 "kernel_ms": round(rm.kernel_ms, 3), "statuses_match_fp64": same,
                 "max_rel_err_vs_fp64_gpu": err_fp64}
        mb.close()
        print(f"# mixed precision: {dt * 1e3:.1f} ms", file=sys.stderr, flush=True)

    plain = None
    graph_launches = batch.graph_launches()
    batch.set_graph_replay(False)  # the plain path launches each solve directly
    if not args.no_plain and args.plain_sample > 0:
        lat = []
        conv_plain = 0
        for i in range(args.plain_sample):
            t0 = time.perf_counter()
            r1 = batch.evaluate(theta[i:i + 1])
            lat.append(time.perf_counter() - t0)
            conv_plain += int(r1.convergence_mask[0])
            if not np.array_equal(r1.predictions[0], last.predictions[i], equal_nan=True):
                raise SystemExit("bench.py: a draw differs alone and in the batch")
        plain = {"value": conv_plain / sum(lat), "unit": "converged solves/s",
                 "sample": f"first {args.plain_sample} draws, one launch each",
                 "latency_ms_p50": round(1e3 * float(np.percentile(lat, 50)), 1)}

    cpu = None
    cpu1 = None
    err = None
    if not args.no_cpu_baseline:
        # the reference's own solver where its harness was built, else the port
        ref = reference_cpu_baseline(prob, theta, _cores())
        rows, cpu = ref if ref is not None else cpu_baseline(prob, theta, _cores())
        errs = []
        for i, o in enumerate(rows):
            if o["status"] != last.status_codes[i]:
                raise SystemExit(f"bench.py: draw {i} status {last.status_codes[i]} vs oracle "
                                 f"{o['status']}")
            if o["status"] in (0, 1):
                p, q = last.predictions[i], o["predictions"]
                errs.append(float(np.max(np.abs(p - q) / np.abs(q))))
        err = max(errs) if errs else None
        cpu["host"] = _host()
        # one worker: the first draw alone
        one = reference_cpu_baseline(prob, theta, 1)
        if one is not None:
            t1 = one[0][0]["wall_seconds"]
            cpu1 = {"value": round(int(one[0][0]["status"] == 0) / t1, 4),
                    "unit": "converged solves/s", "cores": 1, "kind": "reference",
                    "sample": f"draw 0 of the sweep, one process of the reference's solver, "
                              f"{t1:.2f} s"}
        else:
            rows1, cpu1 = cpu_baseline(prob, theta, 1)

    single = None
    if not args.no_single:
        single = _leg("single_solve", single_solve_leg, not args.no_cpu_baseline)

    served = None
    if not args.no_serve:
        served = _leg("served", served_leg, args.serve_requests, args.serve_rate,
                      args.serve_hills, args.serve_plain_check)

    warm = None
    if not args.no_warm:
        warm = _leg("warm_start", warm_leg, 2)

    warm_fd = None
    if not args.no_warm_fd:
        warm_fd = _leg("warm_fd_jacobian", warm_fd_leg, not args.no_cpu_baseline)

    hills = None
    if not args.no_hills:
        hills = _leg("hills", hills_leg, not args.no_hills_full, _cores(),
                     not args.no_cpu_baseline, tuple(args.hill_meshes.split(",")),
                     args.hills_block_size, args.hills_grid_blocks, args.hills_plain_launch)

    calibration = None
    if not args.no_calibration:
        calibration = _leg("calibration", calibration_leg)

    timing = batch.last_timing()
    batch.close()  # every engine is released before the interpreter exits
    line = {
        "metric": "converged SST solves/s (served vs plain) at <=1e-6 rel err; p50/p99 latency",
        "value": round(value, 4),
        "unit": "converged solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(statistics.mean(step_ms), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "64-draw channel sweep: Student-t(4) draws of (a1, beta*), seed "
                               "20261010, 1x128 geometric channel (ratio 1.05), Re_tau 182.088, "
                               "tol 1e-8, PCG pressure / BiCGSTAB transport, 30 probes",
                   "draws": int(len(theta)), "cells": N, "path": "batched persistent kernel"},
        "converged": converged,
        "statuses": {str(s): int(np.sum(last.status_codes == s))
                     for s in np.unique(last.status_codes)},
        "iterations_mean": float(np.mean(last.iteration_counts)),
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 1),
                       "p99": round(float(np.percentile(step_ms, 99)), 1),
                       "note": "batch step latency: every draw of a step completes together"},
        "plain": plain,
        "speedup_vs_plain": round(value / plain["value"], 2) if plain else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_1worker": cpu1,
        "max_rel_err_vs_cpu_baseline": err,
        "mixed": mixed,
        "single_solve": single,
        "served": served,
        "warm_start": warm,
        "warm_fd_jacobian": warm_fd,
        "hills": hills,
        "hill_120x64_converged_solves_per_s": (hills or {}).get("hill_120x64_full", {}).get(
            "converged_per_s"),
        "calibration": calibration,
        "block_size": timing["block_size"],
        "autotune": tune,
        "graph_replay": {"enabled": not args.no_graph, "graph_launches": graph_launches},
        "arena_in_lds": timing["arena_in_lds"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
