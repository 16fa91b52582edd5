"""Benchmark: LPs solved per second by the MI355X restarted-PDHG engine, plus the roofline of
its dominant kernel and a CPU baseline.

Workload (default, "zipf-200k-500k-fp64"): the bandwidth-scale synthetic LP, m = 200,000 rows,
n = 500,000 columns, 10 nonzeros per column at Zipf(1.5)-distributed rows (nnz = 5,000,000;
152 rows hold 88% of the nonzeros), 70% equality / 30% `<=` rows, x0 ~ U[0,1], bounds [0,10],
b = A x0, c = A^T y0 + z (seed 2). Plain FP64.

A step is one whole solve() from the host Model: host preparation (validation, norms, CSR),
upload over PCIe, device preconditioning, power iteration, the iterations, the report and the
release; instance generation is outside the timed region. `value` is LPs per second over the
timed steps. Solved for `value` = relative KKT error (the max of the relative primal, dual and
gap residuals the reference computes) <= 1e-6, the metric's tolerance (engine option
pdhg_stop_rule=relative_kkt). That is NOT the reference's "solved" (meets_request: the same
three <= 1e-6 AND absolute primal <= 1e-7 in the model's units): on this LP the absolute clause
(|b| in the hundreds on the dense rows) is not met within 1e6 iterations, and the engine then
reports iteration_limit, claiming no feasibility. `config.absolute_primal` gives the value
reached. The reference's definition is measured in `reference_rule` on the 1,000 x 3,000 LPs,
where it is reached.

Other fields of the line:
  - `resident`: the same solves of the LP held in HBM (uploaded once; each solve preconditions
    again on the device from the kept values), and whether they end on the same iteration and
    point bit for bit as the whole solve();
  - `reference_rule`: the 1,000 x 3,000 uniform-equality LPs (seeds 1 and 2) solved under the
    reference's own stopping test at 1e-6 (pdhg_stop_at_request), GPU whole solves, beside the
    CPU oracle's full solve of seed 1 on one core and on all cores (A^T y gathered on every
    core, as the reference's threads option does);
  - `mixed`: the same 200k x 500k LP with FP32 iterate deltas and FP64 ticks;
  - `c2`: the 1,000,000 x 2,000,000 Zipf LP (nnz 20M, seed 3) in mixed precision to relative
    KKT 1e-6, then the FP64 polish toward 1e-8 (iteration cap), with the objective error against
    an FP64 solve of the same LP at 1e-6;
  - `batched`: 512 LPs of the batched workload (m log-uniform in [200, 20000], n = 2.5 m, 6
    nonzeros per column) solved in one batch in mixed precision to relative KKT 1e-6, and again
    polished to 1e-8 in FP64 (pdhg_batch_promote), each against plain FP64 single-LP solves of a
    sample of the same LPs at the same tolerance; `speedup` is the ratio of the LPs/s;
  - `adjoint`: 64 blend-style LPs (m = 5,000, n = 12,000) solved toward relative KKT 2e-10 within
    1,500,000 iterations, then the adjoint: forward and backward seconds, and the envelope error
    of L = c.x* (dL/db against y*) over the LPs with a square active system;
  - `roofline`: the dominant kernel's algorithmic bytes over its HIP-event-timed duration;
  - `cpu_threads` (and `c2.cpu_threads`): the oracle on the same LP with one thread and with
    every host core, setup timed alone, then milliseconds per iteration over a fixed budget
    (200 iterations here, 50 on the 1M x 2M LP), beside `cpu_baseline`'s extrapolated rate.

Prints one JSON line. One GPU only: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "sih-26_amd", "python"), os.path.join(ROOT, "oracle")]

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (kind, m, n, per_col, zipf exponent, seed, tolerance, stop rule)
    "zipf-200k-500k-fp64": ("zipf", 200_000, 500_000, 10, 1.5, 2, 1e-6, "relative_kkt"),
    # The 1,000 x 3,000 uniform-equality LP, stopped by the reference's own test at 1e-6
    # (pdhg_stop_at_request: relative residuals and absolute primal <= 1e-7).
    "uniform-1k-3k-fp64": ("uniform", 1_000, 3_000, 5, 1.5, 1, 1e-6, "reference"),
    # The 1,000,000 x 2,000,000 Zipf LP (nnz 20M), for profile runs of its kernels.
    "zipf-1m-2m": ("zipf", 1_000_000, 2_000_000, 10, 1.5, 3, 1e-6, "relative_kkt"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_bytes(m, n, nnz, precision="fp64"):
    """Algorithmic HBM bytes per launch of the two iteration kernels (int32 indices). k_dual
    streams the CSR of Ahat once for both right-hand sides (ext and dx come from one gather of
    the (x, x') pair) and updates y; k_primal streams the CSR of Ahat^T and updates x. Counted
    once per unique element: the matrix (value + index per nonzero, 4 (rows+1) of row pointers),
    the gathered vector, and the epilogue's per-row vectors.
      fp64: values 8 B; x pairs 16 B, y 8 B; per row y, row bounds, y' (32 B); per column x,
            cost and bounds, x' (40 B).
      mixed (the FP32 phase, FP32 deltas around an FP64 restart point): values 4 B; x pairs 8 B,
            y 4 B; per row the FP32 y and y' (8 B), the bounds (16 B) and the FP64 restart point
            and its product yrest, Ahat xrest (16 B), 40 B; per column 48 B likewise."""
    if precision == "mixed":
        k_dual = 8 * nnz + 4 * (m + 1) + 8 * n + m * (4 + 16 + 16 + 4)
        k_primal = 8 * nnz + 4 * (n + 1) + 4 * m + n * (4 + 24 + 16 + 4)
    else:
        k_dual = 12 * nnz + 4 * (m + 1) + 16 * n + m * (8 + 16 + 8)  # y, row bounds, y' write
        k_primal = 12 * nnz + 4 * (n + 1) + 8 * m + n * (8 + 24 + 8)  # x, cost+bounds, x' write
    return k_primal, k_dual


def kernel_roofline(sh, model, precision, spectral_norm, reps, tol=1e-6, stop_rule=1):
    """Both iteration kernels launched standalone on the engine's stream (pdhg_dev_time_kernels:
    HIP events around `reps` launches each) on a handle that has run one graph chunk, with
    their algorithmic bytes (kernel_bytes) and the achieved rate."""
    m, n, nnz = model.dims
    dev = sh.Device(model, precision=precision)
    dev.start(tolerance=tol, spectral_norm=spectral_norm, stop_rule=stop_rule)
    dev.run(1)
    p_ms, d_ms = dev.time_kernels(reps)
    dev.close()
    bp, bd = kernel_bytes(m, n, nnz, precision)
    kp = {"name": "k_primal", "ms": p_ms, "bytes": bp, "gbs": bp / (p_ms * 1e6)}
    kd = {"name": "k_dual", "ms": d_ms, "bytes": bd, "gbs": bd / (d_ms * 1e6)}
    return kp, kd


def traffic_bytes(kernel, args):
    """HBM bytes per launch of `kernel` on this workload, from the committed rocprofv3 PMC
    summary (profiles/pmc_traffic.json: FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    what the counter reads for a copy kernel of known size), or None when not measured."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)
        return rec[args.workload][args.precision][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(arrays, gpu_iterations, tolerance, budget_s=20.0):
    """The oracle (a CPU port of the reference's PDHG, one core) on a bounded sample of the
    same LP: setup alone, then a fixed iteration budget; extrapolated to the iterations the
    GPU needed."""
    import oracle as orc

    P = orc.Problem(arrays)
    t = time.perf_counter()
    r0 = orc.solve(P, tolerance=tolerance, iteration_limit=0)
    setup = time.perf_counter() - t
    iters = 50
    t = time.perf_counter()
    orc.solve(P, tolerance=tolerance, iteration_limit=iters)
    per_iter = max((time.perf_counter() - t - setup) / iters, 1e-12)
    if budget_s > setup + per_iter * iters * 3 and per_iter * 200 < budget_s:
        iters2 = int(min(400, (budget_s - setup) / per_iter))
        t = time.perf_counter()
        orc.solve(P, tolerance=tolerance, iteration_limit=iters2)
        per_iter = max((time.perf_counter() - t - setup) / iters2, 1e-12)
        iters = iters2
    est = setup + per_iter * gpu_iterations
    return {"value": 1.0 / est, "unit": "LPs/s", "cores": 1, "kind": "port",
            "sample": f"oracle on the same LP: setup {setup:.2f}s (scaling + power iteration) "
                      f"timed alone, then {iters} iterations at {per_iter*1e3:.1f} ms/iteration, "
                      f"extrapolated to the GPU's {gpu_iterations} iterations "
                      f"({est:.1f}s per LP); status of the sample run {r0['status']}"}


def cpu_threads(arrays, tolerance, iterations, all_cores):
    """The oracle with one thread and with every host core (only its Âᵀy is parallel, as in the
    reference's threads option) on a fixed iteration budget: setup timed alone, then seconds per
    iteration, for the CPU rows beside the GPU's."""
    import oracle as orc

    P = orc.Problem(arrays)
    out = {"iterations": iterations}
    for label, threads in (("one_core", 1), ("all_cores", max(1, all_cores))):
        orc.set_threads(threads)
        t = time.perf_counter()
        orc.solve(P, tolerance=tolerance, iteration_limit=0)
        setup = time.perf_counter() - t
        t = time.perf_counter()
        r = orc.solve(P, tolerance=tolerance, iteration_limit=iterations)
        per = max((time.perf_counter() - t - setup) / iterations, 1e-12)
        out[label] = {"threads": threads, "setup_seconds": setup, "ms_per_iteration": per * 1e3,
                      "status": r["status"]}
    orc.set_threads(1)
    return out


def _single_sample(sh, models, opts, tol):
    t = time.perf_counter()
    single = [sh.solve(m, **opts) for m in models]
    ts = time.perf_counter() - t
    ok = sum(1 for s in single if s.reached(tol))
    return ok, ts


def batched_comparison(sh, count, polished_sample=4, polished_lps=64, polished_cap=500_000):
    """Mixed-precision batched solves of `count` LPs of the batched workload to relative KKT
    1e-6, against FP64 single-LP solves of a 16-LP sample of the same LPs (every count/16-th),
    whose LPs/s stand for the whole workload. Then the FP64-polished form: the first
    `polished_lps` LPs in one batch, FP32 phase to 1e-6 and pdhg_batch_promote toward 1e-8, under
    an iteration cap (PDHG's relative primal residual levels off near 1e-8 on many of these LPs),
    against FP64 single-LP solves of `polished_sample` of them under the same cap; solved there
    means relative KKT <= 1e-8."""
    import sankhya_hip.workloads as wl

    opts = dict(pdhg_tolerance=1e-6, pdhg_stop_rule="relative_kkt", iteration_limit=1_000_000)
    t = time.perf_counter()
    models = wl.batched(count)
    log(f"[bench] batched workload: {count} LPs, nnz {sum(m.dims[2] for m in models)} "
        f"(generated in {time.perf_counter()-t:.1f}s)")
    t = time.perf_counter()
    sols = sh.solve_batch(models, pdhg_precision="mixed", **opts)
    tb = time.perf_counter() - t
    solved = sum(1 for s in sols if s.reached(1e-6))
    log(f"[bench] batched mixed: {solved}/{count} solved in {tb:.2f}s")
    sample = models[:: max(1, count // 16)]
    single_ok, ts = _single_sample(sh, sample, opts, 1e-6)
    single_rate = single_ok / ts if ts > 0 else 0.0
    log(f"[bench] fp64 single-LP sample: {single_ok}/{len(sample)} in {ts:.2f}s")
    its = sorted(s.iterations for s in sols)
    persistent = [s.stats["schedule"] for s in sols if s.stats["schedule"] >= 0]
    out = {"workload": f"batched-{count}: m log-uniform [200, 20000], n = 2.5 m, 6 nnz/col, "
                       "relative KKT <= 1e-6",
           "value": solved / tb, "unit": "LPs/s", "precision": "mixed", "seconds": tb,
           "tuner": {"cooperative": count - len(persistent),
                     "persistent_by_geometry": [persistent.count(g) for g in range(3)],
                     "seconds_outside_the_loop": (sols[0].solve_seconds - sols[0].stats["loop_seconds"])
                     if sols else None},
           "solved": solved, "lps": count, "max_iterations": its[-1],
           "median_iterations": its[len(its) // 2],
           "fp64_single_value": single_rate, "fp64_single_sample": len(sample),
           "fp64_single_seconds": ts,
           "speedup": (solved / tb) / single_rate if single_rate > 0 else None}
    if polished_sample > 0 and polished_lps > 0:
        popts = dict(opts, pdhg_tolerance=1e-8, iteration_limit=polished_cap)
        pmodels = models[:polished_lps]
        t = time.perf_counter()
        psols = sh.solve_batch(pmodels, pdhg_precision="mixed", **popts)
        tp = time.perf_counter() - t
        psolved = sum(1 for s in psols if s.reached(1e-8))
        fp32 = sorted(s.stats["fp32_iterations"] for s in psols)
        pits = sorted(s.iterations for s in psols)
        log(f"[bench] batched mixed, FP64-polished toward 1e-8: {psolved}/{len(pmodels)} in {tp:.2f}s")
        psample = pmodels[:: max(1, len(pmodels) // polished_sample)]
        pok, t = 0, time.perf_counter()
        for k, md in enumerate(psample):
            s1 = sh.solve(md, **popts)
            pok += 1 if s1.reached(1e-8) else 0
            log(f"[bench] fp64 single LP {k} at 1e-8: {s1.status} kkt {s1.relative_kkt:.1e} "
                f"it={s1.iterations}")
        pts = time.perf_counter() - t
        prate = pok / pts if pts > 0 else 0.0
        log(f"[bench] fp64 single-LP sample at 1e-8: {pok}/{len(psample)} in {pts:.2f}s")
        out["polished_1e8"] = {
            "what": f"the first {len(pmodels)} LPs: FP32 phase to relative KKT 1e-6, then every LP "
                    f"continued in FP64 toward 1e-8, at most {polished_cap} iterations; solved = "
                    "relative KKT <= 1e-8",
            "value": psolved / tp, "unit": "LPs/s", "seconds": tp, "solved": psolved,
            "lps": len(pmodels), "iteration_cap": polished_cap,
            "median_iterations": pits[len(pits) // 2], "max_iterations": pits[-1],
            "median_fp32_iterations": fp32[len(fp32) // 2],
            "fp64_single_value": prate, "fp64_single_sample": len(psample),
            "fp64_single_seconds": pts,
            "speedup": (psolved / tp) / prate if prate > 0 else None}
    return out


def reference_rule_check(sh, seeds=(1, 2), cpu=True):
    """The reference's own definition of solved (meets_request at 1e-6 with
    pdhg_stop_at_request: relative primal, dual, gap <= 1e-6 and absolute primal <= 1e-7) on the
    1,000 x 3,000 uniform-equality LPs: GPU whole solves of each seed, and the CPU oracle's full
    solve of the first seed on one core and on all cores."""
    opts = dict(pdhg_tolerance=1e-6, pdhg_stop_at_request=True, pdhg_polish=False,
                iteration_limit=1_000_000)
    models = [sh.Model.generate("uniform", 1000, 3000, 5, s) for s in seeds]
    for md in models[:1]:
        sh.solve(md, **dict(opts, iteration_limit=1000))  # warm-up (graph capture paths)
    t = time.perf_counter()
    sols = []
    for sd, md in zip(seeds, models):
        sols.append(sh.solve(md, **opts))
        log(f"[bench] reference rule seed {sd}: {sols[-1].status} it={sols[-1].iterations}")
    tg = time.perf_counter() - t
    solved = sum(1 for s in sols if s.status in ("optimal", "feasible"))
    log(f"[bench] reference rule, 1000x3000 seeds {list(seeds)}: {solved}/{len(seeds)} solved in "
        f"{tg:.2f}s on the GPU; iterations {[s.iterations for s in sols]}")
    out = {"workload": "uniform-1k-3k: m = 1,000 equality rows, n = 3,000, 5 nnz/col, seeds "
                       f"{list(seeds)}, the reference's test at 1e-6 (stop at request)",
           "value": solved / tg, "unit": "LPs/s", "seconds": tg, "solved": solved,
           "statuses": [s.status for s in sols], "iterations": [s.iterations for s in sols],
           "objectives": [s.objective for s in sols]}
    if cpu:
        import oracle as orc

        P = orc.Problem(models[0].arrays())
        legs = {}
        for label, threads in (("one_core", 1), ("all_cores", os.cpu_count() or 1)):
            orc.set_threads(threads)
            t = time.perf_counter()
            r = orc.solve(P, tolerance=1e-6, stop_at_request=True, iteration_limit=1_000_000)
            tc = time.perf_counter() - t
            legs[label] = {"value": (1.0 / tc) if r["status"] in ("optimal", "feasible") else 0.0,
                           "unit": "LPs/s", "seconds": tc, "threads": threads, "status": r["status"],
                           "iterations": r["iterations"], "objective": r["objective"]}
            log(f"[bench] oracle seed {seeds[0]} on {threads} thread(s): {r['status']} "
                f"it={r['iterations']} {tc:.2f}s")
        orc.set_threads(1)
        g = sols[0]
        legs["gpu_vs_oracle_seed"] = {
            "status_equal": g.status == legs["one_core"]["status"],
            "objective_rel_diff": abs(g.objective - legs["one_core"]["objective"]) /
                                  max(1.0, abs(legs["one_core"]["objective"]))}
        out["cpu_oracle"] = legs
    return out


def c2_check(sh, iteration_cap=30_000, reps=20, cpu_iterations=50):
    """The 1,000,000 x 2,000,000 Zipf LP (nnz 20M, seed 3): mixed precision, FP32 phase to
    relative KKT 1e-6, then the FP64 polish toward 1e-8 under an iteration cap; and a plain FP64
    solve to 1e-6 whose objective the mixed result is compared with."""
    t = time.perf_counter()
    model = sh.Model.generate("zipf", 1_000_000, 2_000_000, 10, 3)
    log(f"[bench] c2: generated 1M x 2M zipf in {time.perf_counter()-t:.1f}s")
    base = dict(pdhg_stop_rule="relative_kkt")
    t = time.perf_counter()
    a = sh.solve(model, **dict(base, pdhg_tolerance=1e-6, iteration_limit=1_000_000))
    ta = time.perf_counter() - t
    t = time.perf_counter()
    b = sh.solve(model, **dict(base, pdhg_tolerance=1e-6, iteration_limit=1_000_000,
                               pdhg_precision="mixed"))
    tb = time.perf_counter() - t
    t = time.perf_counter()
    c = sh.solve(model, **dict(base, pdhg_tolerance=1e-8, iteration_limit=iteration_cap,
                               pdhg_precision="mixed"))
    tc = time.perf_counter() - t
    rel = lambda s: abs(s.objective - a.objective) / max(1.0, abs(a.objective))
    # Both precisions' iteration kernels on this LP: FP32 halves the matrix values and the
    # gathered pairs, but the FP32-delta epilogue reads the FP64 restart point as well, and a
    # random 8-byte gather costs the same cache line as a 16-byte one.
    kernels = {}
    for prec in ("fp64", "mixed"):
        kp, kd = kernel_roofline(sh, model, prec, a.stats["spectral_norm"], reps)
        kernels[prec] = {k["name"]: {"us": k["ms"] * 1e3, "bytes_per_launch": k["bytes"],
                                     "gbs": k["gbs"], "frac": k["gbs"] / HBM_PEAK_GBS}
                         for k in (kp, kd)}
        log(f"[bench] c2 {prec}: k_primal {kp['ms']*1e3:.1f} us ({kp['gbs']:.0f} GB/s), "
            f"k_dual {kd['ms']*1e3:.1f} us ({kd['gbs']:.0f} GB/s)")
    log(f"[bench] c2: fp64 1e-6 it={a.iterations} {ta:.2f}s; mixed 1e-6 it={b.iterations} {tb:.2f}s "
        f"(rel {rel(b):.1e}); mixed+polish it={c.iterations} (fp32 {c.stats['fp32_iterations']}) "
        f"kkt {c.relative_kkt:.2e} {tc:.2f}s")
    cpu = None
    if cpu_iterations > 0:
        cpu = cpu_threads(model.arrays(), 1e-6, cpu_iterations, host_info()[1])
        log(f"[bench] c2 oracle: {cpu['one_core']['ms_per_iteration']:.1f} ms/iteration on one core, "
            f"{cpu['all_cores']['ms_per_iteration']:.1f} on {cpu['all_cores']['threads']}")
    return {"workload": "zipf-1m-2m: m = 1,000,000, n = 2,000,000, nnz 20M, seed 3",
            "fp64_1e6": {"seconds": ta, "iterations": a.iterations, "relative_kkt": a.relative_kkt,
                         "status": a.status},
            "mixed_1e6": {"seconds": tb, "iterations": b.iterations, "relative_kkt": b.relative_kkt,
                          "objective_rel_error_vs_fp64": rel(b), "status": b.status,
                          "speedup_vs_fp64": ta / tb if tb > 0 else None},
            "kernels": kernels,
            "cpu_threads": cpu,
            "mixed_polished_toward_1e8": {
                "seconds": tc, "iterations": c.iterations, "iteration_cap": iteration_cap,
                "fp32_iterations": c.stats["fp32_iterations"], "relative_kkt": c.relative_kkt,
                "relative_primal": c.stats["final_primal"], "relative_dual": c.stats["final_dual"],
                "relative_gap": c.stats["final_gap"], "reached_1e8": c.reached(1e-8),
                "objective_rel_diff_vs_fp64_1e6": rel(c), "status": c.status}}


def adjoint_check(sh, count, tolerance=2e-10, floor=1e-9, iteration_limit=1_500_000):
    """The differentiable-LP workload: `count` blend LPs (m = 5,000, n = 12,000, seeds 5000+i)
    solved in one batch asking relative KKT <= `tolerance` (within `iteration_limit`), then the
    adjoint (the backward of the torch layer) for L = c.x* and for L = w.x* (w ~ N(0,1)).
    Envelope check of the first, dL/db against the row duals y*, relative to the largest entry
    (and, beside it, in the 2-norm), on every LP whose forward ended at relative KKT <= `floor`
    (whether or not it reached `tolerance` within the limit) and whose active system is square
    (free columns as many as rows with entries: an empty row, 0 = 0, leaves A one rank short);
    a degenerate vertex (fewer free columns than rows: a basic column at its bound) has no
    unique dual, and there the check is that dL/db is an exact dual of the active set,
    A_F^T (dL/db) = c_F. The envelope error is the forward's own dual accuracy (the adjoint
    solves the active system exactly): on this workload it was 1.0e-5 at worst with the forward
    at 1e-9 and 1.8e-6 at 2e-10."""
    import numpy as np
    import scipy.sparse as sps
    import sankhya_hip.workloads as wl

    t = time.perf_counter()
    models = wl.blend(count)
    datas = [m.arrays() for m in models]
    log(f"[bench] adjoint workload: {count} blend LPs (generated in {time.perf_counter()-t:.1f}s)")
    t = time.perf_counter()
    ses = sh.BatchSession(models, pdhg_tolerance=tolerance, pdhg_stop_rule="relative_kkt",
                          iteration_limit=iteration_limit, pdhg_polish=False)
    tf = time.perf_counter() - t
    sols = ses.solutions
    solved = [s.reached(floor) for s in sols]
    reached = sum(1 for s in sols if s.reached(tolerance))
    log(f"[bench] adjoint forward: {reached}/{count} at {tolerance:g}, {sum(solved)}/{count} at "
        f"{floor:g} in {tf:.2f}s")
    t = time.perf_counter()
    gc, gb, its, _ = ses.adjoint([d["cost"] for d in datas], None)
    te = time.perf_counter() - t
    rng = np.random.default_rng(7)
    t = time.perf_counter()
    ses.adjoint([rng.standard_normal(d["n"]) for d in datas], None)
    tw = time.perf_counter() - t
    env_b, env_c, deg_resid, square, degenerate, env_all, env_l2 = 0.0, 0.0, 0.0, 0, 0, [], 0.0
    for k, (d, s) in enumerate(zip(datas, sols)):
        if not solved[k]:
            continue
        x, y = s.col_value, s.row_dual
        lo, hi = d["col_lower"], d["col_upper"]
        slack = np.minimum(np.where(np.abs(lo) < 1e30, x - lo, np.inf),
                           np.where(np.abs(hi) < 1e30, hi - x, np.inf))
        free = np.where(slack > np.abs(s.col_dual))[0]
        env_c = max(env_c, float(np.abs(gc[k]).max()))  # dL/dc - x* for L = c.x*
        if len(free) == int(np.count_nonzero(np.bincount(d["rowidx"], minlength=d["m"]))):
            square += 1
            env_all.append(float(np.abs(gb[k] - y).max() / max(1.0, np.abs(y).max())))
            env_l2 = max(env_l2, float(np.linalg.norm(gb[k] - y) / max(1.0, np.linalg.norm(y))))
            env_b = max(env_b, env_all[-1])
        else:
            degenerate += 1
            # on the columns free beyond doubt (the tie column may read either way)
            sure = np.where(slack > 10.0 * np.abs(s.col_dual))[0]
            A = sps.csc_matrix((d["val"], d["rowidx"], d["colptr"]), shape=(d["m"], d["n"]))
            cf = d["cost"][sure]
            deg_resid = max(deg_resid, float(np.abs(A[:, sure].T @ gb[k] - cf).max() /
                                             max(1.0, np.abs(cf).max())))
    ses.close()
    log(f"[bench] adjoint backward: c.x* {te:.2f}s, w.x* {tw:.2f}s; envelope error on {square} "
        f"square LPs {env_b:.1e}; {degenerate} degenerate (active-system residual {deg_resid:.1e})")
    return {"workload": f"blend-{count}: m = 5000, n = 12000, 40% of columns upper-bounded, "
                        f"forward asked relative KKT <= {tolerance:g} within {iteration_limit:,} "
                        f"iterations, envelope checked where it ended <= {floor:g}; FP64 adjoint",
            "lps": count, "reached_tolerance": reached, "solved": int(sum(solved)),
            "forward_seconds": tf,
            "backward_seconds_objective": te, "backward_seconds_wx": tw,
            "cgls_iterations_median": int(np.median(its[:, 0])),
            "square_lps": square, "degenerate_lps": degenerate,
            "envelope_rel_error_b": env_b,
            "envelope_rel_error_b_median": float(np.median(env_all)) if env_all else None,
            "envelope_rel_error_b_l2": env_l2,
            "envelope_abs_error_c": env_c,
            "degenerate_active_system_residual": deg_resid, "envelope_target": 1e-5}


def host_info():
    """CPU model (lscpu's model name) and logical CPU count (nproc) of this host."""
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except Exception:
        nproc = os.cpu_count()
    return model, nproc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="zipf-200k-500k-fp64", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timing-reps", type=int, default=50)
    ap.add_argument("--iteration-limit", type=int, default=1_000_000,
                    help="iteration limit of the headline solves (profiling runs cap it)")
    ap.add_argument("--batched-lps", type=int, default=512,
                    help="LPs in the batched comparison (0 skips it)")
    ap.add_argument("--polished-sample", type=int, default=4,
                    help="FP64 single-LP sample of the batch's 1e-8 comparison (0 skips that leg)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision solve")
    ap.add_argument("--no-reference-rule", action="store_true", help="skip the reference-rule section")
    ap.add_argument("--no-c2", action="store_true", help="skip the 1M x 2M section")
    ap.add_argument("--adjoint-lps", type=int, default=64,
                    help="blend LPs in the adjoint check (0 skips it)")
    ap.add_argument("--profile-run", action="store_true",
                    help="for rocprofv3: one device handle only, one resident solve of the "
                         "workload and the kernel timing launches, then a short JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not part of this build)"}))
        sys.exit(2)

    import sankhya_hip as sh

    ok, desc = sh.device_description(0)
    if not ok:
        raise SystemExit(f"bench needs a gfx950 device: {desc}")
    kind, m, n, per_col, zexp, seed, tol, rule = WORKLOADS[args.workload]
    stop_req = rule == "reference"
    log(f"[bench] device {desc}; generating {args.workload} (seed {seed})")
    t = time.perf_counter()
    model = sh.Model.generate(kind, m, n, per_col, seed, zexp)
    M, N, NNZ = model.dims
    log(f"[bench] generated m={M} n={N} nnz={NNZ} in {time.perf_counter()-t:.1f}s")
    opts = dict(pdhg_tolerance=tol, pdhg_stop_at_request=stop_req, pdhg_stop_rule=rule,
                iteration_limit=args.iteration_limit, pdhg_precision=args.precision)
    if args.profile_run:
        # One handle for the whole process: rocprofv3's kernel trace has crashed the host process
        # (inside the profiler, during hipGraphLaunch) at the first graph launch of a second
        # handle created after the first was released; with one handle it traces cleanly.
        dev = sh.Device(model, precision=args.precision)
        t = time.perf_counter()
        st, *_ = dev.solve(tolerance=tol, stop_rule=1 if rule == "relative_kkt" else 0,
                           stop_at_request=stop_req, iteration_limit=args.iteration_limit)
        ts = time.perf_counter() - t
        p_ms, d_ms = dev.time_kernels(args.timing_reps)
        dev.close()
        log(f"[bench] profile run: {st.iteration} iterations in {ts:.3f}s; k_primal {p_ms*1e3:.2f} us, "
            f"k_dual {d_ms*1e3:.2f} us")
        print(json.dumps({"mode": "profile-run", "workload": args.workload, "iterations": int(st.iteration),
                          "seconds": ts, "k_primal_us": p_ms * 1e3, "k_dual_us": d_ms * 1e3}), flush=True)
        return

    # ---- headline: whole solve() calls, upload included ----
    for w in range(args.warmup):
        s = sh.solve(model, **opts)
        log(f"[bench] warmup {w}: {s.status} obj={s.objective:.12g} it={s.iterations} "
            f"t={s.solve_seconds:.3f}s")
    times, sols = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        s = sh.solve(model, **opts)
        times.append(time.perf_counter() - ts)
        sols.append(s)
        log(f"[bench] step {k}: {s.status} obj={s.objective:.12g} it={s.iterations} "
            f"restarts={s.stats['restarts']} t={times[-1]:.3f}s (setup {s.stats['setup_seconds']:.3f}s, "
            f"loop {s.stats['loop_seconds']:.3f}s)")
    total = time.perf_counter() - t0
    # Solved: relative KKT <= tol under the relative_kkt rule (no feasibility claimed when the
    # absolute primal clause is missed), or the status's claim under the reference's rule.
    solved = sum(1 for s in sols if s.reached(tol))
    value = solved / total if total > 0 else 0.0
    last = sols[-1]

    # ---- the same LP resident in HBM (secondary) ----
    resident = None
    if args.precision == "fp64" and args.steps > 0:
        rdev = sh.Device(model)
        ropts = dict(tolerance=tol, stop_rule=1 if rule == "relative_kkt" else 0,
                     stop_at_request=stop_req, iteration_limit=args.iteration_limit)
        rsteps = min(args.steps, 5)
        st, *_ = rdev.solve(**ropts)
        rsolved, t1 = 0, time.perf_counter()
        for k in range(rsteps):
            st, x, y, _, _ = rdev.solve(**ropts)
            rsolved += 1 if st.done else 0
        rtotal = time.perf_counter() - t1
        rdev.close()
        same = bool(st.iteration == last.iterations and np.array_equal(x, last.col_value)
                    and np.array_equal(-y, last.row_dual))
        resident = {"value": rsolved / rtotal if rtotal > 0 else 0.0, "unit": "LPs/s",
                    "ms_per_solve": 1e3 * rtotal / max(rsteps, 1), "steps": rsteps,
                    "iterations": int(st.iteration), "identical_to_whole_solve": same,
                    "what": "the LP held in HBM, uploaded once; each solve preconditions again on "
                            "the device from the kept values, nothing uploaded from the host"}
        log(f"[bench] resident: {rsolved}/{rsteps} in {rtotal:.3f}s; same point as the whole "
            f"solve: {same}")

    # ---- roofline of the dominant kernel: standalone launches on the engine's stream, timed
    # with HIP events, on the same uploaded LP ----
    kp, kd = kernel_roofline(sh, model, args.precision, last.stats["spectral_norm"],
                             args.timing_reps, tol, 1 if rule == "relative_kkt" else 0)
    p_ms, d_ms = kp["ms"], kd["ms"]
    dom = kd if d_ms >= p_ms else kp
    log(f"[bench] k_primal {p_ms*1e3:.1f} us {kp['gbs']:.0f} GB/s; k_dual {d_ms*1e3:.1f} us {kd['gbs']:.0f} GB/s")

    mixed = None
    if not args.no_mixed and args.precision == "fp64":
        ts = time.perf_counter()
        sm = sh.solve(model, **dict(opts, pdhg_precision="mixed"))
        tm = time.perf_counter() - ts
        mixed = {"value": (1.0 / tm) if sm.reached(tol) else 0.0,
                 "unit": "LPs/s", "status": sm.status, "iterations": sm.iterations,
                 "fp32_iterations": sm.stats["fp32_iterations"], "seconds": tm,
                 "objective_rel_diff_vs_fp64": abs(sm.objective - last.objective) / max(1.0, abs(last.objective))}
        log(f"[bench] mixed: {sm.status} it={sm.iterations} t={tm:.3f}s")

    reference = None
    if not args.no_reference_rule:
        reference = reference_rule_check(sh, cpu=not args.no_cpu_baseline)

    c2 = None
    if not args.no_c2:
        c2 = c2_check(sh, reps=args.timing_reps, cpu_iterations=0 if args.no_cpu_baseline else 50)

    batched = None
    if args.batched_lps > 0:
        batched = batched_comparison(sh, args.batched_lps, args.polished_sample)

    adjoint = None
    if args.adjoint_lps > 0:
        adjoint = adjoint_check(sh, args.adjoint_lps)

    cpu_model, ncpu = host_info()
    cpu = cpu_rows = None
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU baseline (oracle, one core) on a bounded sample")
        cpu = cpu_baseline(model.arrays(), last.iterations, tol)
        cpu["sample"] += f"; host {cpu_model}, nproc {ncpu}"
        cpu_rows = cpu_threads(model.arrays(), tol, 200, ncpu)
        log(f"[bench] oracle on {args.workload}: {cpu_rows['one_core']['ms_per_iteration']:.1f} "
            f"ms/iteration on one core, {cpu_rows['all_cores']['ms_per_iteration']:.1f} on "
            f"{cpu_rows['all_cores']['threads']}")

    line = {
        "metric": "LPs solved/sec to 1e-6 rel KKT + SpMV HBM GB/s; mixed/batched vs plain FP64",
        "value": value,
        "unit": "LPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (deterministic generator, seeded)",
        "config": {"workload": args.workload, "m": M, "n": N, "nnz": NNZ,
                   "tolerance": tol, "stop_rule": rule, "precision": args.precision,
                   "solved_means": ("relative KKT <= tolerance (pdhg_stop_rule=relative_kkt); not "
                                    "the reference's meets_request, which also needs absolute "
                                    "primal <= 1e-7 (see reference_rule)") if rule == "relative_kkt"
                   else "the reference's meets_request (stop at request)",
                   "step": "one whole solve() from the host Model, upload over PCIe included",
                   "iterations": last.iterations, "status": last.status,
                   "objective": last.objective, "restarts": last.stats["restarts"],
                   "relative_kkt": last.relative_kkt,
                   "absolute_primal": last.stats["final_absolute_primal"]},
        "roofline": {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": traffic_bytes(dom["name"], args),
                     "kernel": dom["name"],
                     "kernel_us": dom["ms"] * 1e3, "bytes_per_launch": dom["bytes"],
                     "other_kernel": {k: v for k, v in (kp if dom is kd else kd).items()}},
        "resident": resident,
        "reference_rule": reference,
        "cpu_baseline": cpu,
        "cpu_threads": cpu_rows,
        "host": {"cpu": cpu_model, "nproc": ncpu},
        "mixed": mixed,
        "c2": c2,
        "batched": batched,
        "adjoint": adjoint,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
