"""Independent acceptance check of an LP answer, restating the LP checks of the reference's
tools/verify_solution.py (:141-160 primal against the claimed status, :190-231 row activity
relative to its terms and activity agreement, :240-253 objective recomputed, :318-355 reduced
costs c - A^T y, :357-404 dual sign conditions, :406-447 complementary slackness, :500-560
strong duality with per-item accounting). Vectorised with numpy; shares no code with the
engine. Test infrastructure.

Inputs: the model arrays (CSC) and a solution dict with status, objective, col_value,
row_dual, col_dual, row_activity (the engine's own, in the model's sense).
"""
from __future__ import annotations

import math

import numpy as np

PRIMAL_TOL = 1e-7
DUAL_TOL = 1e-7
DUALITY_TOL = 1e-9
POINT = {"optimal", "feasible", "unbounded", "iteration_limit", "time_limit", "node_limit",
         "interrupted"}
ASSERTS_FEASIBILITY = {"optimal", "feasible"}


def _inf(v):
    v = np.asarray(v, float)
    return np.where(np.abs(v) >= 1e30, np.copysign(np.inf, v), v)


def verify(d, sol):
    """Returns a list of (check name, ok, detail)."""
    out = []
    status = sol["status"]
    if status not in POINT:
        return [("verdict", True, f"status {status} claims no point")]
    m, n = int(d["m"]), int(d["n"])
    colptr, rowidx, val = np.asarray(d["colptr"]), np.asarray(d["rowidx"]), np.asarray(d["val"], float)
    cols = np.repeat(np.arange(n), np.diff(colptr))
    x = np.asarray(sol["col_value"], float)
    cl, cu = _inf(d["col_lower"]), _inf(d["col_upper"])
    rl, ru = _inf(d["row_lower"]), _inf(d["row_upper"])
    terms = val * x[cols]
    activity = np.bincount(rowidx, weights=terms, minlength=m)
    row_scale = np.ones(m)
    np.maximum.at(row_scale, rowidx, np.abs(terms))
    rviol = np.maximum(np.maximum(rl - activity, activity - ru), 0.0)
    rviol = np.nan_to_num(rviol, nan=0.0, posinf=np.inf)
    worst_rel = float(np.max(rviol / row_scale)) if m else 0.0
    cviol = np.maximum(np.maximum(cl - x, x - cu), 0.0)
    worst_col = float(np.max(cviol / np.maximum(1.0, np.abs(x)))) if n else 0.0
    if status in ASSERTS_FEASIBILITY:
        out.append(("row activity", worst_rel <= PRIMAL_TOL, f"worst relative {worst_rel:.3e}"))
        out.append(("column bounds", worst_col <= PRIMAL_TOL, f"worst relative {worst_col:.3e}"))
    if sol.get("row_activity") is not None and m:
        gap = float(np.max(np.abs(activity - np.asarray(sol["row_activity"], float))))
        out.append(("activity agreement", gap <= 1e-6, f"max gap {gap:.3e}"))
    objective = float(d.get("objective_offset", 0.0)) + float(np.dot(np.asarray(d["cost"], float), x))
    scale = max(1.0, abs(objective))
    out.append(("objective", abs(objective - sol["objective"]) <= 1e-9 * scale,
                f"recomputed {objective:.12e} vs {sol['objective']:.12e}"))
    if status != "optimal":
        return out
    sigma = -1.0 if d.get("maximize", False) else 1.0
    y = sigma * np.asarray(sol["row_dual"], float)
    dd = sigma * np.asarray(sol["col_dual"], float)
    cost = sigma * np.asarray(d["cost"], float)
    yterms = val * y[rowidx]
    aty = np.bincount(cols, weights=yterms, minlength=n)
    cscale = np.maximum(1.0, np.abs(cost))
    np.maximum.at(cscale, cols, np.abs(yterms))
    rc = float(np.max(np.abs(cost - aty - dd) / cscale)) if n else 0.0
    out.append(("reduced costs", rc <= 1e-6, f"worst relative {rc:.3e}"))

    def sign_violation(mult, lo, hi):
        v = np.where((mult > 0) & ~np.isfinite(lo), mult, 0.0)
        v = np.where((mult < 0) & ~np.isfinite(hi), -mult, v)
        return np.where(lo == hi, 0.0, v)

    dcol = float(np.max(sign_violation(dd, cl, cu) / cscale)) if n else 0.0
    out.append(("dual feasibility (columns)", dcol <= DUAL_TOL, f"worst {dcol:.3e}"))
    dual_norm = max(1.0, float(np.max(np.abs(y))) if m else 0.0)
    drow = float(np.max(sign_violation(y, rl, ru) / dual_norm)) if m else 0.0
    out.append(("dual feasibility (rows)", drow <= DUAL_TOL, f"worst {drow:.3e}"))

    def compl(mult, value, lo, hi):
        sl = np.where(np.isfinite(lo), value - lo, np.inf)
        su = np.where(np.isfinite(hi), hi - value, np.inf)
        s = np.minimum(sl, su)
        p = np.where(np.isfinite(s), np.abs(mult) * np.where(np.isfinite(s), s, 0.0), np.abs(mult))
        return np.where(lo == hi, 0.0, p)

    cw = max(float(np.max(compl(y, activity, rl, ru))) if m else 0.0,
             float(np.max(compl(dd, x, cl, cu))) if n else 0.0)
    out.append(("complementary slackness", cw <= 1e-6, f"worst {cw:.3e}"))

    def bound_contribution(mult, lo, hi):
        return np.where(mult > 0, np.where(np.isfinite(lo), mult * np.where(np.isfinite(lo), lo, 0), 0.0),
                        np.where(mult < 0, np.where(np.isfinite(hi), mult * np.where(np.isfinite(hi), hi, 0), 0.0), 0.0))

    def share(mult, value, lo, hi, mscale):
        priced = np.where(mult > 0, lo, hi)
        term = np.abs(mult) * np.where(np.isfinite(priced), np.abs(value - np.where(np.isfinite(priced), priced, 0)), np.abs(value))
        near = np.minimum(np.where(np.isfinite(lo), np.abs(value - np.where(np.isfinite(lo), lo, 0)), np.inf),
                          np.where(np.isfinite(hi), np.abs(value - np.where(np.isfinite(hi), hi, 0)), np.inf))
        small = np.abs(mult) <= DUAL_TOL * mscale
        res = np.where(small, term, np.where(np.isfinite(near), np.minimum(term, np.abs(mult) * np.where(np.isfinite(near), near, 0)), 0.0))
        return np.where(mult == 0, 0.0, res)

    dual_obj_min = float(np.sum(bound_contribution(y, rl, ru)) + np.sum(bound_contribution(dd, cl, cu)))
    accounted = float(np.sum(share(dd, x, cl, cu, cscale)) + np.sum(share(y, activity, rl, ru, dual_norm)))
    dual_objective = sigma * dual_obj_min + float(d.get("objective_offset", 0.0))
    gap = abs(objective - dual_objective)
    out.append(("strong duality", gap <= DUALITY_TOL * scale + accounted,
                f"primal {objective:.12e} dual {dual_objective:.12e} gap {gap:.3e} accounted {accounted:.3e}"))
    return out


def accepted(d, sol):
    checks = verify(d, sol)
    return all(ok for _, ok, _ in checks), checks
