"""tests/verify.py (the acceptance check every GPU solve test applies) against the reference's
own checker, tools/verify_solution.py, whose verdicts tests/golden/make_verifier_golden.py
recorded in this container for the nine Netlib LPs, as solved by the CPU oracle and with three
planted failures (tests/golden/verifier/verdicts.json). For each case the solution file is
written again by the engine's writer from the recorded vectors, read back with tests/solfile.py
and checked by tests/verify.py: the overall verdict and every check the two share by name must
agree with the reference."""
import json
import os

import numpy as np
import pytest

import common as cm
import sankhya_hip as sh
from solfile import read_sol
from verify import verify

FIXTURE = os.path.join(cm.HERE, "golden", "verifier", "verdicts.json")
CASES = json.load(open(FIXTURE))["cases"]


@pytest.mark.parametrize("case", CASES, ids=[f"{c['model']}-{c['variant']}" for c in CASES])
def test_acceptance_check_agrees_with_the_reference_checker(case, tmp_path):
    name = case["model"]
    d = cm.netlib(name)
    mdl = cm.model_of(d)
    mdl.set_names(name.upper(), f"netlib/{name}.mps", d["col_names"], d["row_names"])
    x, yd, dj = (np.array(case[k]) for k in ("col_value", "row_dual", "col_dual"))
    act, q = sh.quality(mdl, x, yd, dj, case["objective"])
    sol = sh.Solution(status=case["status"], objective=float(q[0]), dual_bound=float(case["objective"]),
                      iterations=1, col_value=x, row_dual=yd, col_dual=dj, row_activity=act,
                      algorithm="pdhg-cpu-oracle", message="", solve_seconds=0.0,
                      primal_infeasibility=float(q[1]), dual_infeasibility=float(q[3]),
                      complementarity_violation=float(q[5]), relative_gap=float(q[8]))
    path = tmp_path / "case.sol"
    sh.write_solution(mdl, sol, path)
    f = read_sol(path)
    h = f["header"]
    from_file = dict(status=h["status"], objective=float(h["objective"]),
                     col_value=np.array([f["columns"][c][0] for c in d["col_names"]]),
                     col_dual=np.array([f["columns"][c][1] for c in d["col_names"]]),
                     row_activity=np.array([f["rows"][r][0] for r in d["row_names"]]),
                     row_dual=np.array([f["rows"][r][1] for r in d["row_names"]]))
    ours = {n: ok for n, ok, _ in verify(d, from_file)}
    ref = {n: ok for n, ok in case["checks"]}
    shared = sorted(set(ours) & set(ref))
    differ = [(n, ours[n], ref[n]) for n in shared if ours[n] != ref[n]]
    assert not differ, differ
    assert all(ours.values()) == case["accepted"], (ours, ref)
    if case["variant"] != "as-solved":
        assert not case["accepted"] and not all(ours.values())
    # Every check that failed for the reference is one this restatement also makes.
    assert {n for n, ok in ref.items() if not ok} <= set(ours), (ours, ref)


def test_every_planted_failure_is_caught_by_name():
    caught = {}
    for c in CASES:
        if c["variant"] != "as-solved":
            caught.setdefault(c["variant"], set()).update(n for n, ok in c["checks"] if not ok)
    assert "row activity" in caught["shifted-x"]
    assert "dual feasibility (rows)" in caught["wrong-sign-dual"]
    assert "reduced costs" in caught["broken-reduced-cost"]
