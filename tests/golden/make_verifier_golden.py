"""Writes tests/golden/verifier/verdicts.json: the reference's own acceptance checker
(tools/verify_solution.py, imported from /root/reference/tools) run on solution files that this
engine's writer produced, so that tests/verify.py (the restatement every GPU solve test uses) can
be compared with it check by check.

Run in the build container only (the reference tree never travels to the GPU box):

    python tests/golden/make_verifier_golden.py

Cases: the nine Netlib LPs of tests/golden/netlib solved by the CPU oracle to 1e-8, each as
solved and with three planted failures claimed as optimal: a shifted column value, a sign-flipped
row dual on an inequality row, and a broken reduced cost. The fixture holds the inputs (status,
objective, the three vectors) and the reference's verdict per check; the .sol files are
regenerated from those inputs by the test.
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "sih-26_amd", "python"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)
sys.path.insert(0, "/root/reference/tools")

import oracle as orc  # noqa: E402
import sankhya_hip as sh  # noqa: E402
import verify_solution as vs  # noqa: E402  (the reference's checker, this container only)
from mps import read_mps  # noqa: E402

NETLIB = ["afiro", "sc50a", "sc50b", "blend", "sc105", "adlittle", "stocfor1", "share2b", "israel"]


def model_with_names(name, d):
    import common as cm
    mdl = cm.model_of(d)
    mdl.set_names(name.upper(), f"netlib/{name}.mps", d["col_names"], d["row_names"])
    return mdl


def make_solution(mdl, status, objective, x, yd, dj):
    act, q = sh.quality(mdl, x, yd, dj, objective)
    return sh.Solution(status=status, objective=float(q[0]), dual_bound=float(objective),
                       iterations=1, col_value=np.asarray(x, float), row_dual=np.asarray(yd, float),
                       col_dual=np.asarray(dj, float), row_activity=act, algorithm="pdhg-cpu-oracle",
                       message="", solve_seconds=0.0, primal_infeasibility=float(q[1]),
                       dual_infeasibility=float(q[3]), complementarity_violation=float(q[5]),
                       relative_gap=float(q[8]))


def variants(d, x, yd, dj):
    yield "as-solved", None, x, yd, dj
    lo, hi = np.asarray(d["col_lower"]), np.asarray(d["col_upper"])
    inner = np.flatnonzero((x > lo + 1e-6) & (x < hi - 1e-6) & (np.diff(d["colptr"]) > 0))
    j = int(inner[0]) if len(inner) else 0
    x2 = x.copy()
    x2[j] += 1e-3
    yield "shifted-x", {"column": j, "by": 1e-3}, x2, yd, dj
    rl, ru = np.asarray(d["row_lower"]), np.asarray(d["row_upper"])
    ineq = np.flatnonzero((rl != ru) & (np.abs(yd) > 1e-6))
    if len(ineq):
        i = int(ineq[0])
        y2 = yd.copy()
        y2[i] = -y2[i]
        yield "wrong-sign-dual", {"row": i}, x, y2, dj
    d2 = dj.copy()
    d2[j] += 0.5
    yield "broken-reduced-cost", {"column": j, "by": 0.5}, x, yd, d2


def main():
    cases = []
    with tempfile.TemporaryDirectory() as tmp:
        for name in NETLIB:
            path = os.path.join(ROOT, "tests", "golden", "netlib", f"{name}.mps")
            d = read_mps(path)
            P = orc.Problem(d)
            r = orc.solve(P, tolerance=1e-8, iteration_limit=400000)
            mdl = model_with_names(name, d)
            ref_model = vs.parse_mps(__import__("pathlib").Path(path))
            for variant, planted, x, yd, dj in variants(d, r["col_value"], r["row_dual"], r["col_dual"]):
                status = r["status"] if variant == "as-solved" else "optimal"
                sol = make_solution(mdl, status, r["objective"], x, yd, dj)
                sol_path = os.path.join(tmp, f"{name}-{variant}.sol")
                sh.write_solution(mdl, sol, sol_path)
                report = vs.verify(ref_model, vs.parse_sol(__import__("pathlib").Path(sol_path)),
                                   vs.DEFAULT_PRIMAL_TOL, vs.DEFAULT_DUAL_TOL,
                                   vs.DEFAULT_INTEGER_TOL, vs.DEFAULT_DUALITY_TOL)
                checks = [[n, bool(ok)] for ok, n, _ in report.lines]
                cases.append(dict(model=name, variant=variant, planted=planted, status=status,
                                  objective=r["objective"], col_value=list(map(float, x)),
                                  row_dual=list(map(float, yd)), col_dual=list(map(float, dj)),
                                  accepted=report.failures == 0, checks=checks))
                print(f"{name:9s} {variant:20s} {status:16s} accepted={report.failures == 0} "
                      f"failed={[n for n, ok in checks if not ok]}", flush=True)
    with open(os.path.join(HERE, "verifier", "verdicts.json"), "w") as fh:
        json.dump(dict(checker="tools/verify_solution.py (reference), default tolerances",
                       cases=cases), fh, indent=0)


if __name__ == "__main__":
    main()
