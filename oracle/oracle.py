"""ctypes wrapper of the CPU oracle (liboracle.so) and, when built, of oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker or as the timed CPU baseline. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

STATUS = ["not_solved", "optimal", "infeasible", "unbounded", "infeasible_or_unbounded",
          "feasible", "iteration_limit", "time_limit", "node_limit", "numerical_error",
          "model_error", "interrupted"]

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OrModel(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64),
                ("colptr", C.c_void_p), ("rowidx", C.c_void_p), ("val", C.c_void_p),
                ("cost", C.c_void_p), ("col_lower", C.c_void_p), ("col_upper", C.c_void_p),
                ("row_lower", C.c_void_p), ("row_upper", C.c_void_p),
                ("objective_offset", C.c_double), ("maximize", C.c_int32)]


class OrOptions(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("iteration_limit", C.c_int64),
                ("restart", C.c_int32), ("stop_at_request", C.c_int32),
                ("random_seed", C.c_uint32), ("time_limit", C.c_double), ("stop_rule", C.c_int32)]


class OrResiduals(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("primal", "dual", "gap", "primal_objective",
                                          "dual_objective", "absolute_primal", "absolute_dual",
                                          "gap_as_verified", "complementarity")]

    def as_array(self):
        return np.array([getattr(self, k) for k, _ in self._fields_])


class OrResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int64), ("restarts", C.c_int64),
                ("objective", C.c_double), ("dual_bound", C.c_double),
                ("solve_seconds", C.c_double), ("spectral_norm", C.c_double),
                ("final_residuals", OrResiduals),
                ("primal_infeasibility", C.c_double), ("primal_infeasibility_scaled", C.c_double),
                ("dual_infeasibility", C.c_double), ("dual_infeasibility_scaled", C.c_double),
                ("complementarity_violation", C.c_double), ("absolute_gap", C.c_double),
                ("relative_gap", C.c_double)]


class OrTick(C.Structure):
    _fields_ = [("iteration", C.c_int64), ("objective", C.c_double), ("primal", C.c_double),
                ("dual", C.c_double), ("gap", C.c_double), ("restarted", C.c_int32),
                ("chose_average", C.c_int32)]


class OrMt(C.Structure):
    _fields_ = [("state", C.c_uint32 * 624), ("index", C.c_int)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise OSError(f"{ORACLE_SO} is missing; run make -C oracle")
        L = C.CDLL(ORACLE_SO)
        L.or_spectral_norm.restype = C.c_double
        L.or_spectral_norm.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, C.c_int, C.c_uint32]
        L.or_csc_multiply.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_csc_transpose_multiply.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_build_scaling.argtypes = [C.POINTER(OrModel), _f64p, C.c_int] + [_f64p] * 8 + [C.c_void_p, C.c_void_p]
        L.or_evaluate.argtypes = [C.POINTER(OrModel), _f64p, C.c_double, C.c_double, _f64p, _f64p,
                                  _f64p, _f64p, C.POINTER(OrResiduals)]
        L.or_solve_pdhg.restype = C.c_int
        L.or_solve_pdhg.argtypes = [C.POINTER(OrModel), C.POINTER(OrOptions), _f64p, _f64p, _f64p,
                                    C.POINTER(OrResult), C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_recompute_quality.argtypes = [C.POINTER(OrModel), _f64p, _f64p, _f64p, C.c_double,
                                           _f64p, C.POINTER(OrResult)]
        L.or_mt19937_seed.argtypes = [C.POINTER(OrMt), C.c_uint32]
        L.or_mt19937_next.restype = C.c_uint32
        L.or_mt19937_next.argtypes = [C.POINTER(OrMt)]
        L.or_uniform_real.restype = C.c_double
        L.or_uniform_real.argtypes = [C.POINTER(OrMt), C.c_double, C.c_double]
        _lib = L
    return _lib


def ref() -> Optional[C.CDLL]:
    """The partial reference library, or None when it was not built (no reference tree)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        R = C.CDLL(REF_SO)
        R.ref_build_scaling.restype = C.c_int
        R.ref_build_scaling.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p] + [_f64p] * 6 + [C.c_int] + [_f64p] * 8
        R.ref_spectral_norm.restype = C.c_double
        R.ref_spectral_norm.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, C.c_int, C.c_uint32]
        R.ref_evaluate.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p] + [_f64p] * 6 + \
            [C.c_double, C.c_double, _f64p, _f64p, _f64p, _f64p, _f64p]
        R.ref_recompute_quality.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p] + [_f64p] * 5 + \
            [C.c_double, C.c_int, _f64p, _f64p, _f64p, C.c_double, _f64p, _f64p]
        R.ref_validate.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p] + [_f64p] * 5 + \
            [C.c_char_p, C.c_int]
        _ref = R
    return _ref


class Problem:
    """An LP held as numpy arrays (CSC), with the OrModel view the C functions take."""

    def __init__(self, d: Dict):
        self.m, self.n = int(d["m"]), int(d["n"])
        f = lambda k, t=np.float64: np.ascontiguousarray(d[k], dtype=t)
        self.colptr, self.rowidx = f("colptr", np.int32), f("rowidx", np.int32)
        self.val, self.cost = f("val"), f("cost")
        self.col_lower, self.col_upper = f("col_lower"), f("col_upper")
        self.row_lower, self.row_upper = f("row_lower"), f("row_upper")
        self.offset = float(d.get("objective_offset", 0.0))
        self.maximize = bool(d.get("maximize", False))
        self.nnz = int(self.colptr[-1])
        self.sense = -1.0 if self.maximize else 1.0
        self.min_cost = np.ascontiguousarray(self.sense * self.cost)
        b = np.where(np.abs(self.row_lower) < 1e30, self.row_lower,
                     np.where(np.abs(self.row_upper) < 1e30, self.row_upper, 0.0))
        self.bound_norm = float(np.sqrt(_seq_sumsq(b)))
        self.cost_norm = float(np.sqrt(_seq_sumsq(self.min_cost)))

    def c_model(self) -> OrModel:
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        return OrModel(self.m, self.n, self.nnz, p(self.colptr), p(self.rowidx), p(self.val),
                       p(self.cost), p(self.col_lower), p(self.col_upper), p(self.row_lower),
                       p(self.row_upper), self.offset, 1 if self.maximize else 0)


def _seq_sumsq(v):
    s = 0.0
    for x in v.tolist():
        s += x * x
    return s


def csc_multiply(P: Problem, x, val=None):
    y = np.zeros(max(P.m, 1))
    lib().or_csc_multiply(P.m, P.n, P.colptr, P.rowidx, P.val if val is None else val,
                          np.ascontiguousarray(x, np.float64), y)
    return y[: P.m]


def csc_transpose_multiply(P: Problem, y, val=None):
    x = np.zeros(max(P.n, 1))
    lib().or_csc_transpose_multiply(P.m, P.n, P.colptr, P.rowidx, P.val if val is None else val,
                                    np.ascontiguousarray(y, np.float64), x)
    return x[: P.n]


def build_scaling(P: Problem, ruiz: int = 10) -> Dict[str, np.ndarray]:
    out = {k: np.zeros(max(sz, 1)) for k, sz in (("row", P.m), ("column", P.n), ("val", P.nnz),
                                                   ("cost", P.n), ("col_lower", P.n), ("col_upper", P.n),
                                                   ("row_lower", P.m), ("row_upper", P.m))}
    M = P.c_model()
    lib().or_build_scaling(C.byref(M), P.min_cost, ruiz, out["row"], out["column"], out["val"],
                           out["cost"], out["col_lower"], out["col_upper"], out["row_lower"],
                           out["row_upper"], None, None)
    for k, ln in (("row", P.m), ("column", P.n), ("val", P.nnz), ("cost", P.n), ("col_lower", P.n),
                  ("col_upper", P.n), ("row_lower", P.m), ("row_upper", P.m)):
        out[k] = out[k][:ln]
    return out


def spectral_norm(P: Problem, val, iterations=30, seed=1) -> float:
    return lib().or_spectral_norm(P.m, P.n, P.colptr, P.rowidx, np.ascontiguousarray(val), iterations, seed)


def evaluate(P: Problem, x, y):
    act, red = np.zeros(max(P.m, 1)), np.zeros(max(P.n, 1))
    r = OrResiduals()
    M = P.c_model()
    lib().or_evaluate(C.byref(M), P.min_cost, P.bound_norm, P.cost_norm,
                      np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64),
                      act, red, C.byref(r))
    return act[: P.m], red[: P.n], r.as_array()


def set_threads(threads: int) -> None:
    """Threads of the oracle's A^T y gather (the reference's `threads` option); results do not
    depend on it."""
    lib().or_set_threads(int(threads))


def solve(P: Problem, tolerance=1e-4, iteration_limit=-1, restart=True, stop_at_request=False,
          random_seed=0, time_limit=0.0, max_ticks=0, stop_rule=0):
    x, y, d = np.zeros(max(P.n, 1)), np.zeros(max(P.m, 1)), np.zeros(max(P.n, 1))
    res = OrResult()
    opts = OrOptions(tolerance, iteration_limit, 1 if restart else 0, 1 if stop_at_request else 0,
                     random_seed, time_limit, stop_rule)
    ticks = (OrTick * max(max_ticks, 1))()
    nt = C.c_int64()
    M = P.c_model()
    lib().or_solve_pdhg(C.byref(M), C.byref(opts), x, y, d, C.byref(res),
                        C.cast(ticks, C.c_void_p) if max_ticks else None, max_ticks, C.byref(nt))
    out = {k: getattr(res, k) for k, _ in OrResult._fields_ if k != "final_residuals"}
    out["status"] = STATUS[res.status]
    out["final_residuals"] = res.final_residuals.as_array()
    out["col_value"], out["row_dual"], out["col_dual"] = x[: P.n], y[: P.m], d[: P.n]
    out["ticks"] = [ticks[k] for k in range(min(nt.value, max_ticks))]
    out["num_ticks"] = nt.value
    return out


QUALITY_FIELDS = ("objective", "primal_infeasibility", "primal_infeasibility_scaled",
                  "dual_infeasibility", "dual_infeasibility_scaled", "complementarity_violation",
                  "integrality_violation", "absolute_gap", "relative_gap")


def recompute_quality(P: Problem, x, row_dual, col_dual, dual_bound):
    """or_recompute_quality: (row activity, the nine quality fields in QUALITY_FIELDS order; an LP
    has no integrality, so that field is 0)."""
    act = np.zeros(max(P.m, 1))
    q = OrResult()
    lib().or_recompute_quality(C.byref(P.c_model()), np.ascontiguousarray(x, np.float64),
                               np.ascontiguousarray(row_dual, np.float64),
                               np.ascontiguousarray(col_dual, np.float64), float(dual_bound), act,
                               C.byref(q))
    vals = np.array([q.objective, q.primal_infeasibility, q.primal_infeasibility_scaled,
                     q.dual_infeasibility, q.dual_infeasibility_scaled, q.complementarity_violation,
                     0.0, q.absolute_gap, q.relative_gap])
    return act[: P.m], vals
