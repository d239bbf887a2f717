"""Steady FP64 residual + LU-SGS throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one ``lu_sgs_stage`` (timestep, LSQ gradient + Barth limiter, Lax
flux + chemical source, LU-SGS forward/backward sweeps, primitive update) of
the 400k-triangle nozzle (1000x200 mapped grid, seeded jitter, block-multicolour
cell order), all state resident in device memory.  Prints one JSON line.

roofline: the kernel with the largest share of the step, its algorithmic
bytes per launch (each operand element touched once) over its mean launch
duration, measured with HIP events on the case's stream over a replay of the
timed steps, against the 8 TB/s HBM peak; traffic is that kernel's measured
HBM bytes per launch (rocprofv3 FETCH_SIZE and WRITE_SIZE passes, summarised
in profiles/steady_step_400k_hbm_traffic.md).

cpu_baseline: the reference solver (built from its sources into
oracle/_ref by oracle/ref_build.sh) timed on a bounded sample of the same
workload on one host core; when that binary is absent, the C oracle (a
serial restatement) instead.

Side workloads, each in its own field of the same line (a failing one is
reported there and leaves the headline standing): small_case (the 12k
triangles of the reference's own runs, 2000 steps), serving (continuous
batching of Poisson requests), batched (256 sampled cases, autotuned launch
shape and colouring, against the same cases one at a time), learned (the
FP32 source-term network trained on the device and used behind its gate on
the 1.6M-triangle nozzle); on request viscous and backward_euler.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 cell-updates/s (residual+LU-SGS); served cases/hour + p50/p99 to converge"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell-update of the steady step on triangles (NF/NE = 1.5)
B_CU_TRI = 1256.0 + 160.0 * 1.5


def kernel_bytes(ne: int, nf: int, ng: int) -> dict:
    """Algorithmic (compulsory) bytes per launch of each kernel: every array
    element the kernel must read or write, once; FP64 = 8 B, int32 = 4 B,
    s = 56 B is one 7-variable state."""
    s = 56
    return {
        # PA, stencil, wdis, AMLS, limiter offsets, GradC; ghost normal/type, GPA
        "gradient_limiter": (s + 12 + 24 + 48 + 48 + 2 * s) * ne + (16 + 4 + s) * ng,
        # PA+GradC of every cell; nl, nr, VECX, flux-point offsets, F, alpaF
        "flux_face": 3 * s * ne + (8 + 16 + 32 + s + 8) * nf,
        # cface, PA, vol, hWA, step, dd; F, alpaF
        "residual_source": (12 + s + 8 + s + 16) * ne + (s + 8) * nf,
        # sched, cface, dd, hWA rw, PA of neighbours; VECX, nl, nr, alpaF
        "lusgs_forward": (4 + 12 + 8 + 2 * s + s) * ne + 32 * nf,
        "lusgs_backward": (4 + 12 + 8 + 2 * s + s) * ne + 32 * nf,
        # WA rw, hWA, T, PA
        "update": (2 * s + s + 8 + s) * ne,
    }


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "steady_step_400k_hbm_traffic.md")


def measured_traffic() -> dict:
    """HBM bytes per launch of each kernel of the 1000x200 workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in the profiles
    folder (each counter scaled by what it reads for a plain copy)."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return json.loads(line.split(":", 1)[1])
    except OSError:
        pass
    return {}


def host_cores() -> dict:
    """What the host offers this process: nproc, the CPUs this process may
    run on, and the share used for all-core runs (at most 16 per GPU)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "affinity": aff, "share": max(1, min(16, aff))}


def cpu_baseline_all_cores(case, nsteps: int, procs: int) -> dict:
    """The reference binary on every core of the host share at once: procs
    independent runs of the same case (the reference's own multi-core mode is
    MPI domain decomposition, which needs mpmetis and is not built), started
    together; aggregate cell-updates/s = procs x cells x steps / the slowest
    run's step-loop wall time."""
    ref = os.path.join(ROOT, "oracle", "_ref", "rc_ref_dump")
    if not os.path.exists(ref):
        return {"error": "reference binary not built"}
    from rocketcombustor_amd import nozzle
    tmp = tempfile.mkdtemp(prefix="rcbench_all_")
    try:
        dirs = []
        for k in range(procs):
            d = os.path.join(tmp, "c%d" % k)
            nozzle.write_reference_case(d, case, nstep_max=nsteps)
            os.makedirs(os.path.join(d, "out"))
            dirs.append(d)
        t0 = time.perf_counter()
        ps = [subprocess.Popen([ref, str(nsteps), "0", "out", "1"], cwd=d, stdout=subprocess.PIPE,
                               stderr=subprocess.DEVNULL, text=True) for d in dirs]
        walls = []
        for p in ps:
            out, _ = p.communicate(timeout=900)
            line = [ln for ln in out.splitlines() if ln.startswith("steps")]
            if p.returncode != 0 or not line:
                return {"error": "a reference run failed"}
            walls.append(float(line[-1].split()[3]))
        elapsed = time.perf_counter() - t0
        ne = case.mesh.ne
        return {"value": procs * ne * nsteps / max(walls), "unit": "cell-updates/s", "cores": procs,
                "kind": "reference",
                "sample": "%d concurrent runs of %d steps of the same %d-cell case, reference Fortran (amdflang -O3, "
                          "1 rank each); slowest step loop %.2f s, process wall %.2f s"
                          % (procs, nsteps, ne, max(walls), elapsed)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline(case, nsteps: int) -> dict:
    ref = os.path.join(ROOT, "oracle", "_ref", "rc_ref_dump")
    ne = case.mesh.ne
    if os.path.exists(ref):
        from rocketcombustor_amd import nozzle
        tmp = tempfile.mkdtemp(prefix="rcbench_")
        try:
            nozzle.write_reference_case(tmp, case, nstep_max=nsteps)
            os.makedirs(os.path.join(tmp, "out"))
            r = subprocess.run([ref, str(nsteps), "0", "out", "1"], cwd=tmp, capture_output=True, text=True,
                               timeout=600)
            line = [ln for ln in r.stdout.splitlines() if ln.startswith("steps")]
            stage = [ln for ln in r.stdout.splitlines() if ln.startswith("stage_s")]
            if r.returncode == 0 and line:
                wall = float(line[-1].split()[3])
                out = {"value": ne * nsteps / wall, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                       "sample": "%d steps of the same %d-cell case, reference Fortran (amdflang -O3, 1 rank), "
                                 "step loop wall time %.2f s" % (nsteps, ne, wall)}
                if stage:
                    t = [float(x) for x in stage[-1].split()[2:6]]
                    out["ms_per_step_by_call"] = {k: 1e3 * v / nsteps for k, v in
                                                  zip(("timestep", "reconstruction", "flux_sourceterm",
                                                       "lu_sgs_steady"), t)}
                return out
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    from oracle import pyoracle
    o = pyoracle.OracleCase(case)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.stage()
    wall = time.perf_counter() - t0
    return {"value": ne * nsteps / wall, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the same %d-cell case, serial C restatement" % (nsteps, ne)}


def small_run(nx: int = 150, ny: int = 40, nsteps: int = 2000, cpu_steps: int = 200, cpu: bool = True) -> dict:
    """The reference's own CPU-sized case: the 12k-triangle nozzle (150x40),
    2000 steady steps from the initial field, graph-replayed on the device
    (small meshes are latency-bound: us per step is the figure), and the
    reference binary on the first cpu_steps of the same case."""
    from rocketcombustor_amd import native, nozzle
    from rocketcombustor_amd.solver import SteadyCase
    case = nozzle.nozzle_case(nx, ny)
    perm, _ = native.order_multicolor(case.mesh.cells)
    case = case.permuted(perm)
    with SteadyCase(case, device=0) as g:
        g.run_steady(20, history=False)                 # warm-up, graph capture
        g.set_primitive(case.pa0)                       # back to the initial field
        t0 = time.perf_counter()
        hist = g.run_steady(nsteps, every=nsteps)
        wall = time.perf_counter() - t0
        out = {"workload": "steady nozzle %dx%d, %d triangles, block-4 order, %d steps from the initial field"
                           % (nx, ny, g.ne, nsteps),
               "cells": g.ne, "steps": nsteps, "seconds": wall, "us_per_step": 1e6 * wall / nsteps,
               "cell_updates_per_s": g.ne * nsteps / wall,
               "residual_drop_rho": float(hist[-1][0] / hist[0][0])}
    if cpu:
        c = cpu_baseline(case, cpu_steps)
        c["seconds_for_%d_steps_at_that_rate" % nsteps] = case.mesh.ne * nsteps / c["value"]
        out["cpu_baseline"] = c
        out["speedup_time_to_solution"] = c["seconds_for_%d_steps_at_that_rate" % nsteps] / wall
    return out


def converge_run(case, max_steps: int, every: int = 100, ref_ms_per_step=None) -> dict:
    """The steady run of one case from its initial field, up to max_steps:
    the steps, device seconds and the reference's seconds (at its measured
    rate) to each order of drop of RoRes (the species part of
    step_residual, what the reference monitors, src/residual.f90:79).  A
    drop to 1e-8 is not reached (the Barth limiter and the threshold
    kinetics slow the iteration to a crawl, on the reference too).
    ``still_falling`` says whether the last fifth of the run still went 5%
    below the rest; only when it did not is a plateau reported: the first
    report within 10% of the lowest value of the run."""
    from rocketcombustor_amd.solver import SteadyCase
    from rocketcombustor_amd import native
    out = {"max_steps": max_steps, "report_every": every}
    with SteadyCase(case, device=0) as g:
        t0 = time.perf_counter()
        err = None
        rows = []
        done = 0
        try:
            while done < max_steps:
                k = min(2000, max_steps - done)
                h = g.run_steady(k, every=every)
                # rows of steps done+1 (first) and done+every, done+2*every, ...
                rows.extend(h if done == 0 else h[1:])
                done += k
        except native.RcError as e:
            err = str(e)
        wall = time.perf_counter() - t0
    hist = np.array(rows)
    steps = np.r_[1, np.arange(every, every * (len(hist) - 1) + 1, every)][:len(hist)]
    ro = hist[:, 4:7].sum(axis=1)
    rho = hist[:, 0]

    def plateau(v):
        return int(steps[np.argmax(v <= 1.1 * v.min())])

    cut = max(1, (4 * len(ro)) // 5)

    ms = 1e3 * wall / max(done, 1)
    falling = bool(ro[cut:].min() < 0.95 * ro[:cut].min() or rho[cut:].min() < 0.95 * rho[:cut].min())
    to_drop = {"%g" % d: (int(steps[np.argmax(ro <= d * ro[0])]) if np.any(ro <= d * ro[0]) else None)
               for d in (1e-1, 1e-2, 1e-3)}
    out.update({"steps_run": int(done), "error": err, "ms_per_step": ms,
                "roress_drop": float(ro.min() / ro[0]), "rho_drop": float(rho.min() / rho[0]),
                "still_falling": falling, "steps_to_drop": to_drop,
                "seconds_to_drop": {k: (v * ms * 1e-3 if v else None) for k, v in to_drop.items()}})
    if ref_ms_per_step:
        out["reference_seconds_to_drop_at_its_rate"] = {k: (v * ref_ms_per_step * 1e-3 if v else None)
                                                       for k, v in to_drop.items()}
        out["speedup_time_to_drop"] = ref_ms_per_step / ms
    if not falling:
        ps = max(plateau(ro), plateau(rho))
        out["plateau_step"] = ps
        out["seconds_to_plateau"] = ps * ms * 1e-3
        if ref_ms_per_step:
            out["reference_seconds_to_plateau_at_its_rate"] = ps * ref_ms_per_step * 1e-3
    return out


SERVE_RATE = 20.0     # requests per second
SERVE_ORDERS = 2.0    # converged: RoRes dropped two orders from its first step
SERVE_ROUND = 200     # steps between convergence checks


def serving_run(n: int, plain: bool) -> dict:
    """Continuous-batching serving of n seeded requests (Poisson arrivals,
    5k-400k cells, parameters of the batched family), and optionally the
    same requests served one at a time through the single-case path."""
    from rocketcombustor_amd import serving
    reqs = serving.poisson_requests(n, rate=SERVE_RATE)
    srv = serving.Server(round_steps=SERVE_ROUND, orders=SERVE_ORDERS, max_steps=20000, workers=8)
    rep = srv.run(reqs, time_limit=600.0)
    if rep.get("cell_updates_per_s"):
        rep["step_hbm_frac_bcu"] = B_CU_TRI * rep["cell_updates_per_s"] / (HBM_PEAK_GBS * 1e9)
    out = {"requests": n, "rate_req_per_s": SERVE_RATE, "cells": "log-uniform 5e3-4e5",
           "converged_when": "RoRes <= 1e-2 x its step-1 value (max 20000 steps)", "batched": rep,
           "hbm_frac_definition": "%.0f B per cell-update (1256 + 160 NF/NE at NF/NE = 1.5) x cell-updates/s "
                                  "/ 8 TB/s" % B_CU_TRI}
    if plain:
        seq = serving.serve_sequential(serving.poisson_requests(n, rate=SERVE_RATE), round_steps=SERVE_ROUND,
                                       orders=SERVE_ORDERS, max_steps=20000, time_limit=900.0)
        if seq.get("cell_updates_per_s"):
            seq["step_hbm_frac_bcu"] = B_CU_TRI * seq["cell_updates_per_s"] / (HBM_PEAK_GBS * 1e9)
        out["plain"] = seq
        if seq["cases_per_hour"] > 0:
            out["speedup_cases_per_hour"] = rep["cases_per_hour"] / seq["cases_per_hour"]
    return out


def batched_run(n: int, steps: int, tune: bool = True) -> dict:
    """n seeded independent cases (cells log-uniform 5k-80k, chamber
    pressure, T0, equivalence and throat ratio sampled): the autotuner picks
    the colouring granularity and the launch shape on these cases, then
    `steps` steady steps of all of them as one batch (graph replay, one
    launch per kernel for all cases) with the tuned setting and with the
    default shape, then the same steps case by case through the single-case
    path."""
    import concurrent.futures as cf
    from rocketcombustor_amd import autotune, nozzle, serving
    from rocketcombustor_amd.solver import SteadyBatch
    specs = nozzle.sample_cases(n, seed=3)
    tuned = None
    if tune:
        t0 = time.perf_counter()
        tuned = autotune.tune(specs, steps=10, warmup=2, workers=8)
        tuned["seconds"] = time.perf_counter() - t0
    block = tuned["block"] if tuned else 4
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=8) as pool:
        cases = list(pool.map(lambda sp: serving.build_case(sp, block=block), specs))
    t_build = time.perf_counter() - t0
    cells = sum(g.ne for g in cases)
    nbytes = sum(g.device_bytes for g in cases)
    with SteadyBatch() as b:
        for g in cases:
            b.add(g)
        b.run_steady(5, history=False)          # warm-up (graph capture)
        t0 = time.perf_counter()
        b.run_steady(steps, history=False)
        t_default = time.perf_counter() - t0
        if tuned:
            b.set_launch(tuned["cell_threads"], tuned["run_threads"], tuned["xcd"])
            b.run_steady(2, history=False)
            t0 = time.perf_counter()
            b.run_steady(steps, history=False)
            t_batch = time.perf_counter() - t0
        else:
            t_batch = t_default
        nlev = b.info()["levels"]
        for g in cases:
            b.remove(g)
    for g in cases:
        g.run_steady(2, history=False)          # warm-up of each case's own graph
    t0 = time.perf_counter()
    for g in cases:
        g.run_steady(steps, history=False)
    t_plain = time.perf_counter() - t0
    for g in cases:
        g.close()
    out = {"cases": n, "cells": cells, "device_bytes": nbytes, "steps": steps, "levels": nlev,
           "ordering_block": block, "host_build_s": t_build,
           "batched_cell_updates_per_s": cells * steps / t_batch, "batched_ms_per_step": 1e3 * t_batch / steps,
           "batched_default_shape_ms_per_step": 1e3 * t_default / steps,
           "plain_cell_updates_per_s": cells * steps / t_plain, "plain_ms_per_step_all": 1e3 * t_plain / steps,
           "speedup": t_plain / t_batch,
           "batched_step_hbm_frac_bcu": B_CU_TRI * cells * steps / t_batch / (HBM_PEAK_GBS * 1e9)}
    if tuned:
        trials = tuned.pop("trials")
        tuned["trials_fastest5"] = sorted(trials, key=lambda t: t["ms_per_step"])[:5]
        tuned["trials_slowest"] = max(trials, key=lambda t: t["ms_per_step"])
        tuned["ntrials"] = len(trials)
        out["autotune"] = tuned
    return out


def learned_run(train_steps: int, batch: int = 8192, lr: float = 2e-3, nsteps: int = 50, tol: float = 0.05,
                rtol: float = 1e-3) -> dict:
    """The learned-chemistry workload: ~10M states of the 12k and 400k
    steady runs (snapshots with the device's exact FP64 rates), the MLP
    (6 -> 3 x 128 tanh -> 3, FP32) trained on the device, its error on
    held-out states, then the 1.6M-triangle nozzle stepped with exact
    kinetics, with the surrogate behind the trust gate (training range, mass
    balance and a trust region calibrated on held-out states: no exact rate
    formed), behind the plausibility gate (range and mass balance only) and
    behind the verified gate (also within rtol of the exact FP64 rates)."""
    import numpy as np
    from rocketcombustor_amd import native, nozzle, surrogate
    from rocketcombustor_amd.solver import SteadyCase

    def ordered(nx, ny):
        c = nozzle.nozzle_case(nx, ny)
        perm, _ = native.order_multicolor(c.mesh.cells)
        return c.permuted(perm)

    def rel_err(w, we, tf):
        return np.abs(w - we) / np.maximum(np.abs(we), 1e-3 * tf.out_scale)

    t0 = time.perf_counter()
    snaps = [1] + list(range(50, 1201, 50))
    raw, omega = surrogate.sample_states([ordered(150, 40), ordered(1000, 200)], steps=snaps,
                                         per_snapshot=400000, seed=5)
    t_sample = time.perf_counter() - t0
    rng = np.random.default_rng(5)
    # held out: 100k states to state the error, 400k to calibrate the trust
    # region of the gate that forms no exact rate
    out_idx = rng.choice(len(raw), size=500000, replace=False)
    hold, calib = out_idx[:100000], out_idx[100000:]
    keep = np.ones(len(raw), dtype=bool)
    keep[out_idx] = False
    tf = surrogate.Transform.fit(raw[keep], omega[keep], knee=1e-3)
    m = surrogate.MLP(seed=5)
    m.set_data(tf.x(raw[keep]), tf.y(omega[keep]))
    t1 = time.perf_counter()
    losses = surrogate.train_schedule(m, train_steps, batch=batch, lr=lr, seed=5, phases=6, decay=0.4)
    t_train = time.perf_counter() - t1
    flops_step = 3 * 2 * batch * sum(fi * fo for fi, fo in m.shapes)   # forward + 2x backward
    w_hold = tf.omega(m.forward(tf.x(raw[hold])))
    rel = rel_err(w_hold, omega[hold], tf)
    out = {"states": int(len(raw)), "sample_s": t_sample,
           "network": "6 -> 3 x 128 tanh -> 3, FP32, outputs omega_k = 1e-3 s_k sinh(y_k)",
           "train": {"rows": int(keep.sum()), "steps": train_steps, "batch": batch, "lr": lr,
                     "schedule": "Adam, 6 equal phases, lr x0.4 per phase", "loss_per_phase": losses.tolist(),
                     "seconds": t_train, "tflops": flops_step * train_steps / t_train / 1e12},
           "holdout": {"rows": int(len(hold)), "rel_err_median": float(np.median(rel)),
                       "rel_err_p90": float(np.percentile(rel, 90)), "rel_err_p99": float(np.percentile(rel, 99)),
                       "rel_err_definition": "|w - w_exact| / max(|w_exact|, 1e-3 s_k), per species",
                       "pass_rtol_1e-3": float(surrogate.gate(tf, raw[hold], w_hold, omega[hold], 1e9, 1e-3).mean())}}
    w_cal = tf.omega(m.forward(tf.x(raw[calib])))
    trust = surrogate.TrustRegion.calibrate(tf, raw[calib], w_cal, omega[calib], dims=(48, 16, 48), tau=2.5e-4,
                                            nmin=10)
    hold_ok = surrogate.gate(tf, raw[hold], w_hold, omega[hold], tol, 0.0, trust)
    out["trust_region"] = {"features": "T, log p, Y_1", "dims": list(trust.dims), "tau": trust.tau,
                           "nmin": trust.nmin, "calibration_rows": int(len(calib)),
                           "accepted_grid_cells": int(trust.accept.sum()),
                           "holdout_accepted_fraction": float(hold_ok.mean()),
                           "holdout_accepted_rel_err_max": float(rel[hold_ok].max(axis=1).max()) if hold_ok.any()
                           else None}
    big_case = ordered(2000, 400)
    res, fields = {}, {}
    for mode in ("exact", "trust_gate", "plausibility_gate", "verified_gate"):
        with SteadyCase(big_case, device=0) as g:
            if mode != "exact":
                g.set_surrogate(m, tf, tol=tol, rtol=rtol if mode == "verified_gate" else 0.0,
                                trust=trust if mode == "trust_gate" else None)
            g.run_steady(5, history=False)
            t2 = time.perf_counter()
            g.run_steady(nsteps, history=False)
            wall = time.perf_counter() - t2
            res[mode] = {"ms_per_step": 1000.0 * wall / nsteps, "cell_updates_per_s": g.ne * nsteps / wall}
            pa = g.primitive()
            fields[mode] = pa
            if mode != "exact":
                acc, ev = g.surrogate_stats()
                res[mode]["fallback_fraction"] = 1.0 - acc / max(ev, 1)
                # the final state: the gate restated on the host splits the
                # cells; the accepted cells' rates against the exact ones
                w_exact = g.derived()["production_rate"]
                r = surrogate.raw_inputs(pa)
                w_net = tf.omega(m.forward(tf.x(r)))
                ok = surrogate.gate(tf, r, w_net, w_exact, tol, rtol if mode == "verified_gate" else 0.0,
                                    trust if mode == "trust_gate" else None)
                e = rel_err(w_net, w_exact, tf)
                res[mode]["final_state"] = {
                    "accepted_fraction": float(ok.mean()),
                    "accepted_rel_err_median": float(np.median(e[ok])) if ok.any() else None,
                    "accepted_rel_err_max": float(e[ok].max()) if ok.any() else None}
                dT = np.abs(pa[:, 3] - fields["exact"][:, 3]) / fields["exact"][:, 3]
                res[mode]["T_rel_diff_vs_exact_after_%d_steps" % (nsteps + 5)] = float(dT.max())
                res[mode]["network_ns_per_cell"] = (res[mode]["ms_per_step"] - res["exact"]["ms_per_step"]) * 1e6 / g.ne
            res["cells"] = g.ne
    out["nozzle_1p6M"] = res
    out["gate"] = {"tol": tol, "rtol": rtol}
    m.close()
    return out


def viscous_run(case, nsteps: int, cpu: bool = True) -> dict:
    """Steady steps of the same case on the viscous path (IfViscous: the
    viscous spectral radius in TimeStep, flux_viscous in flux_sourceterm),
    graph-replayed like the inviscid ones, after 5 untimed steps."""
    import dataclasses
    from rocketcombustor_amd import deck
    from rocketcombustor_amd.solver import SteadyCase
    c = dataclasses.replace(case, info=dataclasses.replace(case.info, viscous=True),
                            transport=deck.default_transport())
    with SteadyCase(c, device=0) as g:
        g.run_steady(5, history=False)
        t0 = time.perf_counter()
        g.run_steady(nsteps, history=False)
        wall = time.perf_counter() - t0
        res = g.residual()
        assert all(x == x for x in res), "non-finite residual"
        out = {"workload": "same case with IfViscous, %d steps" % nsteps, "ms_per_step": 1000.0 * wall / nsteps,
               "cell_updates_per_s": g.ne * nsteps / wall}
    if cpu:
        ref = cpu_baseline(c, 3)
        if ref is not None:
            ref["sample"] = ref["sample"].replace("steps of the same", "viscous steps of the same")
            out["cpu_baseline"] = ref
    return out


def backward_euler_run(case, nsteps: int, dt: float = 1e-7, inner: int = 4, cpu: bool = True) -> dict:
    """Dual-time backward Euler (implicit_time_method=3) on the same case:
    ``nsteps`` physical steps of exactly ``inner`` inner iterations (relTol
    0), each inner iteration TimeStep, reconstruction, flux_sourceterm, the
    point-implicit LU-SGS (a 7x7 block solve per cell and sweep) and the
    inner-iteration residual, after one untimed warm-up step."""
    from rocketcombustor_amd.solver import SteadyCase
    with SteadyCase(case, device=0) as g:
        g.run_unsteady(1, dt=dt, innerloop=inner, reltol=0.0, a_p=0.0, f_p=0.0)
        t0 = time.perf_counter()
        rows = g.run_unsteady(nsteps, dt=dt, innerloop=inner, reltol=0.0, a_p=0.0, f_p=0.0)
        wall = time.perf_counter() - t0
        n_inner = len(rows)
        assert all(x == x for x in rows[-1]), "non-finite residual"
        out = {"workload": "same case, dt %g, %d steps x %d inner iterations" % (dt, nsteps, inner),
               "inner_iterations": n_inner, "ms_per_inner_iteration": 1000.0 * wall / n_inner,
               "cell_inner_iterations_per_s": g.ne * n_inner / wall}
    if cpu:
        out["cpu_baseline"] = backward_euler_cpu(case, dt)
    return out


def backward_euler_cpu(case, dt: float, inner: int = 2) -> dict:
    """The reference's backward-Euler stage on one host core: one step of
    ``inner`` inner iterations of the same case (oracle/_ref/rc_ref_dump)."""
    import dataclasses
    ref = os.path.join(ROOT, "oracle", "_ref", "rc_ref_dump")
    if not os.path.exists(ref):
        return None
    from rocketcombustor_amd import nozzle
    c = dataclasses.replace(case, info=dataclasses.replace(case.info, steady=False, implicit_time_method=3,
                                                           innerloop=inner, reltol=0.0, a_p=0.0, f_p=0.0,
                                                           nstep_max=1, totaltime=dt))
    tmp = tempfile.mkdtemp(prefix="rcbench_be_")
    try:
        nozzle.write_reference_case(tmp, c, nstep_max=1)
        os.makedirs(os.path.join(tmp, "out"))
        r = subprocess.run([ref, "1", "0", "out", "1"], cwd=tmp, capture_output=True, text=True, timeout=900)
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("steps")]
        if r.returncode != 0 or not line:
            return None
        wall = float(line[-1].split()[3])
        ne = case.mesh.ne
        return {"value": ne * inner / wall, "unit": "cell-inner-iterations/s", "cores": 1, "kind": "reference",
                "sample": "1 step x %d inner iterations of the same %d-cell case, reference Fortran (amdflang "
                          "-O3, MKL dgesvx, 1 rank), wall time %.2f s" % (inner, ne, wall)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def copy_calibration(mib: int) -> int:
    """Three device-to-device copies of `mib` MiB through the HIP runtime (its
    copy kernel), as the reference point for the HBM byte counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(3):
            if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:    # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return n


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=1000)
    ap.add_argument("--ny", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-all-cores", dest="cpu_all_cores", action="store_false",
                    help="skip the all-core CPU figure (the reference on every core of the host share at once)")
    ap.add_argument("--serve", type=int, default=24, metavar="N",
                    help="requests of the serving run (Poisson 20/s, 5k-400k cells); 0 skips it")
    ap.add_argument("--serve-plain", type=int, default=1,
                    help="also serve the same requests one case at a time (plain path) for the speedup")
    ap.add_argument("--batched", type=int, default=256, metavar="N",
                    help="also autotune and time N independent sampled cases (5k-80k cells) stepped as one "
                         "batch against the same cases stepped one at a time; 0 skips it")
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--backward-euler", type=int, default=0, metavar="N",
                    help="also time N dual-time backward-Euler steps (point-implicit LU-SGS, 4 inner "
                         "iterations each) of the same 400k-cell case")
    ap.add_argument("--viscous", type=int, default=0, metavar="K",
                    help="also time K steady steps of the same case on the viscous path (IfViscous, the "
                         "shipped transport tables)")
    ap.add_argument("--no-small", dest="small", action="store_false",
                    help="skip the 12k-triangle case (2000 steps on the device, the reference binary on a sample)")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="skip the learned-chemistry workload (by default: sample states from the 12k and 400k "
                         "cases, train the source-term MLP on the device, then step the 1.6M-triangle nozzle with "
                         "the gated surrogate and with exact kinetics)")
    ap.add_argument("--learned-steps", type=int, default=100000, help="training steps of --learned")
    ap.add_argument("--converge-steps", type=int, default=30000, metavar="N",
                    help="also run the same case from its initial field for N steps and report the steps and "
                         "time to each order of residual drop (and the plateau, if it flattens); 0 skips it")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device three times (a plain copy kernel of "
                         "known size, to scale the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one case on one GPU; --gpus %d is not supported "
              "(the multi-GPU design is a domain decomposition, not implemented here)" % args.gpus)
        return 2

    from rocketcombustor_amd import native, nozzle
    from rocketcombustor_amd.solver import SteadyCase

    case = nozzle.nozzle_case(args.nx, args.ny)
    perm, ncolors = native.order_multicolor(case.mesh.cells)
    case = case.permuted(perm)
    g = SteadyCase(case, device=0)
    ne, nf, ng = g.ne, g.nf, g.ng
    g.run_steady(args.warmup, history=False)

    # timed region: K steps replayed as the captured step graph (the
    # library's fast path); run_steady synchronises before returning
    t0 = time.perf_counter()
    g.run_steady(args.steps, history=False)
    wall = time.perf_counter() - t0
    # the same K steps launched eagerly with HIP events recorded between the
    # kernels on the case's stream: per-kernel launch durations
    g.set_timing(True)
    t1 = time.perf_counter()
    g.run_steady(args.steps, history=False)
    wall_eager = time.perf_counter() - t1
    times = g.timing()
    g.set_timing(False)
    res = g.residual()
    assert all(x == x for x in res), "non-finite residual"

    value = ne * args.steps / wall
    ms_step = 1000.0 * wall / args.steps
    kb = kernel_bytes(ne, nf, ng)
    # the step forms the residual inside the forward sweep kernels (its own
    # timing slot stays empty): their bytes are counted there
    kb["lusgs_forward"] += kb.pop("residual_source")
    kernels = {k: times[k] for k in g.TIMING_KEYS if k != "residual_source"}
    # timing slots of the sweeps: the forward levels below the highest, the
    # highest level, the other backward levels.  The highest level's
    # backward sweep runs in its forward launch: it reads nothing that launch
    # has not read and writes its cells' increments once, so it adds no
    # compulsory bytes there; bytes are split by the levels' cell counts
    top_cells = int(g.level_cells()[-1])
    top_fused = os.environ.get("RC_TOPFUSE", "1") != "0"
    low = (ne - top_cells) / ne
    fwd, bwd = kb.pop("lusgs_forward"), kb.pop("lusgs_backward")
    kb["lusgs_forward"] = fwd * low
    kb["lusgs_top"] = fwd * (1 - low)
    kb["lusgs_backward"] = bwd * low if top_fused else bwd
    launches = {k: 1 for k in kernels}
    launches["lusgs_forward"] = max(g.nlev - 1, 1)
    launches["lusgs_backward"] = max(g.nlev - 1, 1) if top_fused else g.nlev
    kernels = {k: v for k, v in kernels.items() if v > 0}
    dom = max(kernels, key=kernels.get)
    # per launch: algorithmic bytes / mean launch duration
    ach = (kb[dom] / launches[dom]) / (kernels[dom] * 1e-3 / launches[dom]) / 1e9
    # the committed counter summary measured an earlier build (five sweep
    # levels, other launch sizes): its bytes per launch are not this build's
    traffic = None
    b_cu = 1256.0 + 160.0 * nf / ne          # whole-step algorithmic bytes per cell-update
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "steady reacting nozzle, %d triangles (%dx%d mapped grid, seeded jitter), "
                               "block-multicolour order (blocks of 4, %d colours), Lax+LSQ+Barth+LU-SGS, CFL 5" %
                               (ne, args.nx, args.ny, ncolors),
                   "cells": ne, "faces": nf, "ghosts": ng, "lusgs_levels": g.nlev,
                   "level_cells": [int(c) for c in g.level_cells()],
                   "highest_level_backward_in_forward_launch": top_fused},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_note": "no counter pass of this build: rocprofv3 stopped at the first dispatch with "
                                     "this library; %s holds an earlier build's per-launch bytes"
                                     % os.path.relpath(TRAFFIC_PROFILE, ROOT),
                     "algorithmic_bytes_per_launch": kb[dom] / launches[dom], "launches_per_step": launches[dom]},
        "kernel_timing": "HIP events between kernels, eager replay of the same %d steps (%.4f ms/step)"
                         % (args.steps, 1000.0 * wall_eager / args.steps),
        "kernel_ms_per_step": {k: round(v, 5) for k, v in kernels.items()},
        "kernel_gbs": {k: round(kb[k] / (kernels[k] * 1e-3) / 1e9, 1) for k in kernels if kernels[k] > 0},
        "step_hbm_frac_bcu": b_cu * value / (HBM_PEAK_GBS * 1e9),
    }
    g.close()
    if args.copy_calibration > 0:
        out["copy_calibration_bytes"] = copy_calibration(args.copy_calibration)

    def leg(name, fn, *a, **kw):
        # a failing side workload is reported in its field; the headline stands
        try:
            out[name] = fn(*a, **kw)
        except Exception as e:  # noqa: BLE001
            out[name] = {"error": "%s: %s" % (type(e).__name__, e)}
            print("bench.py: %s failed: %s" % (name, e), file=sys.stderr, flush=True)

    if args.small:
        leg("small_case", small_run, cpu=not args.no_cpu_baseline)
    if args.viscous > 0:
        leg("viscous", viscous_run, case, args.viscous, cpu=not args.no_cpu_baseline)
    if args.learned:
        leg("learned", learned_run, args.learned_steps)
    if args.backward_euler > 0:
        leg("backward_euler", backward_euler_run, case, args.backward_euler, cpu=not args.no_cpu_baseline)
    if args.serve > 0:
        leg("serving", serving_run, args.serve, args.serve_plain)
    if args.batched > 0:
        leg("batched", batched_run, args.batched, args.batched_steps)
    if args.converge_steps > 0:
        # the reference's per-step time on this case, measured on a sample
        ref_ms = None
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = cpu_baseline(case, args.cpu_steps)
            ref_ms = 1e3 * ne / out["cpu_baseline"]["value"]
        leg("converge_400k", converge_run, case, args.converge_steps, ref_ms_per_step=ref_ms)
    if not args.no_cpu_baseline:
        if "cpu_baseline" not in out:
            out["cpu_baseline"] = cpu_baseline(case, args.cpu_steps)
        hc = host_cores()
        out["cpu_baseline"].update(hc)
        if args.cpu_all_cores:
            try:
                out["cpu_baseline"]["all_cores"] = cpu_baseline_all_cores(case, args.cpu_steps, hc["share"])
            except Exception as e:  # noqa: BLE001
                out["cpu_baseline"]["all_cores"] = {"error": "%s: %s" % (type(e).__name__, e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
