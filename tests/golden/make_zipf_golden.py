"""Writes tests/golden/zipf_oracle.json: the CPU oracle's full solves of row-side Zipf LPs (the
structure of the 200,000 x 500,000 and 1,000,000 x 2,000,000 workloads: dense rows of A on the
row side, 70% equality and 30% <= rows), which take minutes on one core, so the GPU tests
compare against the recorded results instead of running the oracle.

    python tests/golden/make_zipf_golden.py            # every case
    python tests/golden/make_zipf_golden.py 200000x500000   # cases whose key starts so

Records are merged by key into the existing file, so cases can be rerun one at a time.

Each record: the generator call, the reference's stopping test (tolerance, stop_at_request),
status, objective, iterations, restarts, final residuals and the bound norm, and a sample of
the tick trace (iteration, relative primal, dual, gap, restarted) every `trace_every` ticks.

Perturbed records (key suffix -ulp<j><+|->): the same solve with cost[j] moved by one ulp up or
down. PDHG's path is sensitive to rounding, and on these LPs so is its stopping iteration: on
the 2,000 x 5,000 LP the unperturbed oracle meets the standard at 1,227,840 iterations, while a
one-ulp change of one cost entry moves that past 2,000,000. The perturbed records show which
outcomes rounding alone can produce, and so what an engine that rounds differently may end with.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "sih-26_amd", "python"), os.path.join(ROOT, "oracle")]

import oracle as orc  # noqa: E402
import sankhya_hip as sh  # noqa: E402

BASE = [
    # (m, n, per column, seed, tolerance, stop_at_request, iteration limit, stop rule)
    # stop rule 0: the reference's test; 1: relative KKT alone (the engine's
    # pdhg_stop_rule=relative_kkt, which the bandwidth-scale LP is measured with).
    (2000, 5000, 10, 3, 1e-6, True, 2_000_000, 0),
    (2000, 5000, 10, 3, 1e-8, False, 2_000_000, 0),
    (20000, 50000, 10, 3, 1e-6, True, 2_000_000, 0),
    (200000, 500000, 10, 2, 1e-6, False, 1_000_000, 1),
]
# One-ulp moves of one cost entry: (index or -1 for the largest |cost|, direction).
ULP = [(-1, 1), (17, -1), (4242, 1)]
CASES = [c + (None,) for c in BASE] + [c + (u,) for c in BASE[:3] for u in ULP]


def key(case):
    m, n, per_col, seed, tol, sar, limit, rule, ulp = case
    k = f"{m}x{n}-s{seed}-{tol:g}-{'request' if sar else 'standard'}-{'kkt' if rule else 'reference'}"
    return k if ulp is None else k + f"-ulp{'max' if ulp[0] < 0 else ulp[0]}{'+' if ulp[1] > 0 else '-'}"
TRACE_EVERY = 100


def run(case):
    m, n, per_col, seed, tol, sar, limit, rule, ulp = case
    mdl = sh.Model.generate("zipf", m, n, per_col, seed)
    d = mdl.arrays()
    if ulp is not None:
        c = np.array(d["cost"], dtype=np.float64)
        j = int(np.argmax(np.abs(c))) if ulp[0] < 0 else ulp[0]
        c[j] = np.nextafter(c[j], np.inf if ulp[1] > 0 else -np.inf)
        d = dict(d, cost=c)
    P = orc.Problem(d)
    t = time.time()
    r = orc.solve(P, tolerance=tol, stop_at_request=sar, iteration_limit=limit, max_ticks=1 << 20,
                  stop_rule=rule)
    seconds = time.time() - t
    b = np.where(np.abs(d["row_lower"]) < 1e30, d["row_lower"],
                 np.where(np.abs(d["row_upper"]) < 1e30, d["row_upper"], 0.0))
    ticks = r["ticks"]
    rec = dict(key=key(case), base_key=key(case[:-1] + (None,)), ulp=ulp, stop_rule=rule, kind="zipf", m=m, n=n, per_col=per_col, seed=seed, tolerance=tol,
               stop_at_request=sar, iteration_limit=limit, status=r["status"],
               objective=r["objective"], iterations=r["iterations"], restarts=r["restarts"],
               final_residuals=r["final_residuals"].tolist(),
               bound_norm=float(np.sqrt(np.sum(b * b))), seconds=round(seconds, 1),
               trace_every=TRACE_EVERY, ticks=len(ticks),
               trace=[[t.iteration, t.primal, t.dual, t.gap, t.restarted]
                      for t in ticks[::TRACE_EVERY]])
    print(f"{key(case)}: {r['status']} "
          f"{r['objective']:.15g} in {r['iterations']} iterations ({seconds:.0f}s)", flush=True)
    return rec


def write(fresh):
    """Merges the records in `fresh` by key into the fixture, in the order of CASES."""
    path = os.path.join(HERE, "zipf_oracle.json")
    old = json.load(open(path))["cases"] if os.path.exists(path) else []
    merged = {c.get("key"): c for c in old}
    merged.update({c["key"]: c for c in fresh})
    order = [key(c) for c in CASES]
    cases = [merged[k] for k in order if k in merged]
    with open(path, "w") as fh:
        json.dump(dict(oracle="oracle/pdhg_oracle.c or_solve_pdhg (src/pdhg/pdhg.cpp:76-568)",
                       cases=cases), fh)


def main():
    from multiprocessing import Pool
    todo = [c for c in CASES if not sys.argv[1:] or any(key(c).startswith(a) for a in sys.argv[1:])]
    todo.sort(key=lambda c: -c[0] * c[1])  # the longest solves first
    with Pool(min(len(todo), 6)) as pool:  # one core per case; each record is written when done
        for rec in pool.imap_unordered(run, todo):
            write([rec])


if __name__ == "__main__":
    main()
