"""Shared test helpers: instances, known answers, and the GPU-presence check."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import oracle as orc
import sankhya_hip as sh
import sankhya_hip.workloads as wl
from mps import read_mps

HERE = os.path.dirname(os.path.abspath(__file__))
NETLIB = os.path.join(HERE, "golden", "netlib")
INF = math.inf


def gpu_or_skip():
    ok, desc = sh.device_description(0)
    if not ok:
        pytest.skip(f"no gfx950 device: {desc}")
    return desc


def netlib(name: str):
    return read_mps(os.path.join(NETLIB, f"{name}.mps"))


def netlib_optima():
    with open(os.path.join(NETLIB, "reference_optima.json")) as fh:
        return json.load(fh)


def model_of(d) -> sh.Model:
    return sh.Model.from_csc(d["m"], d["n"], d["colptr"], d["rowidx"], d["val"], d["cost"],
                             d["col_lower"], d["col_upper"], d["row_lower"], d["row_upper"],
                             d.get("objective_offset", 0.0), d.get("maximize", False))


def generated(kind, m, n, per_col, seed, param=1.5):
    mdl = sh.Model.generate(kind, m, n, per_col, seed, param)
    return mdl, mdl.arrays()


def c0(seed: int):
    """The 1,000 x 3,000 uniform-equality instance, 5 nonzeros per column."""
    return generated("uniform", 1000, 3000, 5, seed)


def dense_lp(rows, lower, upper, cost, col_upper=None, maximize=False, offset=0.0):
    """Same construction as make_lp in the reference's tests/unit/test_pdhg.cpp:66-89."""
    A = np.asarray(rows, dtype=float)
    m, n = A.shape
    colptr, rowidx, val = [0], [], []
    for j in range(n):
        for i in range(m):
            if A[i, j] != 0.0:
                rowidx.append(i)
                val.append(A[i, j])
        colptr.append(len(rowidx))
    return dict(m=m, n=n, colptr=np.array(colptr, np.int32), rowidx=np.array(rowidx, np.int32),
                val=np.array(val), cost=np.array(cost, float), col_lower=np.zeros(n),
                col_upper=np.array(col_upper if col_upper is not None else [INF] * n, float),
                row_lower=np.array(lower, float), row_upper=np.array(upper, float),
                objective_offset=offset, maximize=maximize)


def blend_guard():
    """reference tests/unit/test_pdhg.cpp:451-472 (make_blend_lp)."""
    d = dense_lp([[1.0, 1.0, 1.0], [0.30, 0.45, 0.38], [0.80, -0.86, -0.22]], [90.0, 40.0, -INF],
                 [120.0, 40.0, 0.0], [2.40, 1.60, 1.64], maximize=True)
    d["col_lower"] = np.array([10.0, 0.0, 0.0])
    d["col_upper"] = np.array([INF, 45.0, 60.0])
    return d


# (name, lp dict, expected objective, tolerance) from test_pdhg.cpp:95-146,474-488
KNOWN_ANSWERS = [
    ("MinimisesASum", lambda: dense_lp([[1, 1]], [2], [INF], [1, 1]), 2.0, 1e-6),
    ("HonoursAnEqualityRow", lambda: dense_lp([[1, 1]], [4], [4], [1, 2]), 4.0, 1e-6),
    ("HonoursARangeRow", lambda: dense_lp([[1]], [2], [6], [1]), 2.0, 1e-6),
    ("HonoursAColumnUpperBound", lambda: dense_lp([[1]], [-INF], [100], [-1], col_upper=[7]), -7.0, 1e-6),
    ("ReportsTheObjectiveInTheOriginalSense",
     lambda: dense_lp([[1, 0], [0, 2], [3, 2]], [-INF] * 3, [4, 12, 18], [3, 5], maximize=True), 36.0, 1e-5),
    ("CarriesTheObjectiveOffset", lambda: dense_lp([[1]], [3], [INF], [1], offset=10.0), 13.0, 1e-6),
]


def kkt_instances():
    """The 120 LPs of the reference's simplex-agreement test (tests/unit/test_pdhg.cpp:178-222,
    generator tests/oracles/lp_generator.cpp:122-184, seed 20260904), as the fixture written by
    tests/golden/kkt_instances.cpp: (lp dict, planted optimum). min c'x, A x >= b, 0 <= x <= u
    (tests/oracles/rational_simplex.cpp:252-280)."""
    with open(os.path.join(HERE, "golden", "kkt_20260904.json")) as fh:
        raw = json.load(fh)
    out = []
    for lp in raw:
        ub = [INF if u < 0 else float(u) for u in lp["upper"]]
        out.append((dense_lp(lp["a"], lp["b"], [INF] * lp["rows"], lp["c"], col_upper=ub),
                    float(lp["objective"])))
    return out


def large_rhs_lp():
    """min x1 + x2, x1 + x2 >= 1e5: relative and absolute residuals five orders of magnitude
    apart (tests/unit/test_pdhg.cpp:224-292)."""
    return dense_lp([[1.0, 1.0]], [1.0e5], [INF], [1.0, 1.0])


PRIMAL_FEASIBILITY = 1e-7  # include/sankhya/tolerances.hpp kPrimalFeasibility


def sparse_lp(m, n, cols, cost, col_lower, col_upper, row_lower, row_upper):
    """An LP from {column: [(row, value), ...]}, min sense."""
    colptr, rowidx, val = [0], [], []
    for j in range(n):
        for i, a in cols.get(j, []):
            rowidx.append(i)
            val.append(a)
        colptr.append(len(rowidx))
    return dict(m=m, n=n, colptr=np.array(colptr, np.int32), rowidx=np.array(rowidx, np.int32),
                val=np.array(val, float), cost=np.array(cost, float),
                col_lower=np.array(col_lower, float), col_upper=np.array(col_upper, float),
                row_lower=np.array(row_lower, float), row_upper=np.array(row_upper, float),
                objective_offset=0.0, maximize=False)


# Degenerate shapes: (name, lp, expected status, expected objective or None). The optima are
# by hand; PDHG never reports infeasibility (src/pdhg/pdhg.cpp:486-567), so an infeasible LP
# ends at the iteration limit.
DEGENERATE = [
    # min x0 - 2 x1 + x2/2 over a box: x = (0, 5, -1)
    ("no_rows", sparse_lp(0, 3, {}, [1, -2, 0.5], [0, 0, -1], [4, 5, 3], [], []), "optimal", -10.5),
    # -1 <= 0 <= 1 and 0 = 0
    ("rows_without_entries", sparse_lp(2, 2, {}, [1, 1], [0, 0], [3, 3], [-1, 0], [1, 0]),
     "optimal", 0.0),
    # column 1 has no entries: x1 at its lower bound 1; x0 + x2 >= 4 at x0 = 4
    ("empty_column", sparse_lp(1, 3, {0: [(0, 1.0)], 2: [(0, 1.0)]}, [1, 2, 3], [0, 1, 0],
                               [10, 10, 10], [4], [INF]), "optimal", 6.0),
    # x0 fixed at 2; x0 + x1 >= 3 and x0 - x1 <= 5: x1 = 1. (A free row, both bounds infinite,
    # is left out: its complementarity term is |y| * inf (src/pdhg/pdhg_evaluate.cpp:103-114),
    # so the stopping test holds only at a tick where that row's y is exactly zero.)
    ("fixed_column", sparse_lp(2, 2, {0: [(0, 1.0), (1, 1.0)], 1: [(0, 1.0), (1, -1.0)]},
                               [1, 1], [2, 0], [2, 10], [3, -INF], [INF, 5]),
     "optimal", 3.0),
    # x0 <= 1 and x0 >= 5
    ("infeasible", sparse_lp(1, 1, {0: [(0, 1.0)]}, [1], [0], [1], [5], [INF]), "iteration_limit", None),
]


def transposed_zipf(m: int, n: int, seed: int):
    """The transpose of the m x n Zipf matrix (10 per column) as an n x m LP: its densest columns
    are nearly full, so the column side of the work plan has long-row segments. Bounds
    0 <= x <= 10, -1 <= Ax <= 1 (x = 0 is feasible), cost uniform in [-1, 1] (seed 5)."""
    import scipy.sparse as sp
    d = generated("zipf", m, n, 10, seed)[1]
    B = sp.csc_matrix((d["val"], d["rowidx"], d["colptr"]), shape=(d["m"], d["n"])).T.tocsc()
    B.sort_indices()
    rows, cols = B.shape
    rng = np.random.default_rng(5)
    lp = dict(m=rows, n=cols, colptr=B.indptr.astype(np.int32), rowidx=B.indices.astype(np.int32),
              val=B.data.copy(), cost=rng.uniform(-1, 1, cols), col_lower=np.zeros(cols),
              col_upper=np.full(cols, 10.0), row_lower=np.full(rows, -1.0),
              row_upper=np.full(rows, 1.0), objective_offset=0.0, maximize=False)
    return model_of(lp), lp


def c3_sizes(count: int, seed: int = 1000):
    """Row counts of the batched workload (sankhya_hip.workloads.batched_sizes)."""
    return wl.batched_sizes(count, seed)


def c3_batch(count: int, seed: int = 1000):
    """`count` LPs of the batched workload (sankhya_hip.workloads.batched)."""
    return wl.batched(count, seed)


def simplex_result(d):
    """The LP solved to a vertex by HiGHS's dual simplex (scipy.optimize.linprog): an FP64
    simplex oracle independent of PDHG. Rows become equalities or one-sided inequalities."""
    import scipy.optimize as so
    import scipy.sparse as sps
    A = sps.csc_matrix((d["val"], d["rowidx"], d["colptr"]), shape=(d["m"], d["n"]))
    rl, ru = np.asarray(d["row_lower"], float), np.asarray(d["row_upper"], float)
    fl, fu = np.abs(rl) < 1e30, np.abs(ru) < 1e30
    eq = fl & fu & (rl == ru)
    ub_hi = fu & ~eq
    ub_lo = fl & ~eq
    A_ub = sps.vstack([A[ub_hi], -A[ub_lo]]).tocsc() if (ub_hi.any() or ub_lo.any()) else None
    b_ub = np.concatenate([ru[ub_hi], -rl[ub_lo]]) if A_ub is not None else None
    cl, cu = np.asarray(d["col_lower"], float), np.asarray(d["col_upper"], float)
    bounds = np.column_stack([np.where(np.abs(cl) < 1e30, cl, -np.inf), np.where(np.abs(cu) < 1e30, cu, np.inf)])
    sign = -1.0 if d.get("maximize") else 1.0
    r = so.linprog(sign * np.asarray(d["cost"], float), A_ub=A_ub, b_ub=b_ub,
                   A_eq=A[eq] if eq.any() else None, b_eq=rl[eq] if eq.any() else None,
                   bounds=bounds, method="highs-ds")
    assert r.status == 0, r.message
    return r, eq


def simplex_objective(d) -> float:
    r, _ = simplex_result(d)
    sign = -1.0 if d.get("maximize") else 1.0
    return sign * r.fun + float(d.get("objective_offset", 0.0))
