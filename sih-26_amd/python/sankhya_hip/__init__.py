"""Python bindings of the MI355X restarted-PDHG engine (ctypes over include/sankhya_pdhg.h and
include/pdhg_hip.h).

Mirrors the shape of the reference's ctypes bindings (bindings/python/sankhya/_library.py and
__init__.py:485-699): a Model built from arrays, `solve()` with options as keyword arguments,
and a Solution with the reference's field names. The shared library is built in-tree
(sih-26_amd/lib/libsankhya_hip.so); importing this module without it raises, and nothing here
falls back to a CPU solver.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "lib", "libsankhya_hip.so"))

STATUS_NAMES = [
    "not_solved", "optimal", "infeasible", "unbounded", "infeasible_or_unbounded", "feasible",
    "iteration_limit", "time_limit", "node_limit", "numerical_error", "model_error",
    "interrupted",
]
PRECISION = {"fp64": 0, "mixed": 1}

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("iterations", C.c_int64), ("attempts", C.c_int64), ("restarts", C.c_int64),
        ("ticks", C.c_int64), ("device_bytes", C.c_int64), ("fp32_iterations", C.c_int64),
        ("schedule", C.c_int32),
        ("objective", C.c_double), ("dual_bound", C.c_double), ("solve_seconds", C.c_double),
        ("setup_seconds", C.c_double), ("loop_seconds", C.c_double),
        ("spectral_norm", C.c_double),
        ("primal_infeasibility", C.c_double), ("dual_infeasibility", C.c_double),
        ("complementarity_violation", C.c_double), ("relative_gap", C.c_double),
        ("final_primal", C.c_double), ("final_dual", C.c_double), ("final_gap", C.c_double),
        ("final_absolute_primal", C.c_double),
        ("algorithm", C.c_char * 32), ("message", C.c_char * 1024),
    ]


class DevState(C.Structure):
    _fields_ = [
        ("iteration", C.c_int64), ("attempts", C.c_int64), ("restarts", C.c_int64),
        ("ticks", C.c_int64), ("done", C.c_int32), ("limit_hit", C.c_int32),
        ("halted", C.c_int32), ("eta", C.c_double), ("omega", C.c_double),
        ("best_worst", C.c_double), ("best_primal", C.c_double), ("best_dual", C.c_double),
        ("best_dual_objective", C.c_double),
    ]


class DevParams(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double), ("iteration_limit", C.c_int64), ("restart", C.c_int32),
        ("stop_at_request", C.c_int32), ("spectral_norm", C.c_double),
        ("chunk_attempts", C.c_int32), ("stop_rule", C.c_int32),
    ]


class Tick(C.Structure):
    _fields_ = [
        ("iteration", C.c_int64), ("primal_objective", C.c_double), ("primal", C.c_double),
        ("dual", C.c_double), ("gap", C.c_double), ("restarted", C.c_int32),
        ("chose_average", C.c_int32),
    ]


_lib: Optional[C.CDLL] = None


def load_library() -> C.CDLL:
    """Load lib/libsankhya_hip.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.sankhya_model_new.restype = vp
    lib.sankhya_model_new.argtypes = [C.c_int32, C.c_int32, C.c_int64, _i32p, _i32p, _f64p, _f64p,
                                      _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int32]
    lib.sankhya_model_generate.restype = vp
    lib.sankhya_model_generate.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                           C.c_uint64]
    lib.sankhya_model_dims.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int64)]
    lib.sankhya_model_export.argtypes = [vp, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                         C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    lib.sankhya_model_free.argtypes = [vp]
    lib.sankhya_pdhg_solve.restype = C.c_int
    lib.sankhya_pdhg_solve.argtypes = [vp, C.c_char_p, vp, vp, vp, vp, C.POINTER(Result)]
    lib.sankhya_pdhg_solve_batch.restype = C.c_int
    lib.sankhya_pdhg_solve_batch.argtypes = [C.POINTER(vp), C.c_int32, C.c_char_p, vp, vp, vp, vp,
                                             C.POINTER(Result)]
    lib.sankhya_batch_solve.restype = vp
    lib.sankhya_batch_solve.argtypes = [C.POINTER(vp), C.c_int32, C.c_char_p, vp, vp, vp, vp,
                                        C.POINTER(Result)]
    lib.sankhya_batch_adjoint.restype = C.c_int
    lib.sankhya_batch_adjoint.argtypes = [vp, _f64p, _f64p, C.c_double, C.c_int, _f64p, _f64p,
                                          np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
                                          _f64p]
    lib.sankhya_batch_free.argtypes = [vp]
    lib.sankhya_batch_adjoint_masks.argtypes = [vp, _f64p, _f64p]
    lib.sankhya_model_set_names.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p),
                                            C.POINTER(C.c_char_p)]
    lib.sankhya_model_name_of.argtypes = [vp, C.c_int32, C.c_int32]
    lib.sankhya_model_name_of.restype = C.c_char_p
    lib.sankhya_model_col_is_integer.argtypes = [vp, C.c_int32]
    lib.sankhya_model_create.restype = vp
    lib.sankhya_model_read.argtypes = [vp, C.c_char_p]
    lib.sankhya_last_error.restype = C.c_char_p
    lib.sankhya_write_solution.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(Result), vp, vp,
                                           vp, vp, C.c_char_p, C.c_int32]
    lib.sankhya_pdhg_scaling.argtypes = [vp, C.c_int, C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p,
                                         _f64p, _f64p, _f64p]
    lib.sankhya_pdhg_device_open.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp)]
    lib.pdhg_dev_probe.argtypes = [C.c_int, C.c_char_p, C.c_int]
    lib.pdhg_dev_destroy.argtypes = [vp]
    lib.pdhg_dev_last_error.restype = C.c_char_p
    lib.pdhg_dev_last_error.argtypes = [vp]
    lib.pdhg_dev_spmv.argtypes = [vp, C.c_int, _f64p, _f64p]
    lib.pdhg_dev_spectral_norm.argtypes = [vp, _f64p, C.c_int, C.POINTER(C.c_double)]
    lib.pdhg_dev_evaluate.argtypes = [vp, _f64p, _f64p, _f64p, _f64p, _f64p]
    lib.pdhg_dev_start.argtypes = [vp, C.POINTER(DevParams)]
    lib.pdhg_dev_run.argtypes = [vp, C.c_int, C.POINTER(DevState)]
    lib.pdhg_dev_promote.argtypes = [vp, C.POINTER(DevParams)]
    lib.pdhg_dev_ticks.argtypes = [vp, C.c_int64, C.POINTER(Tick), C.c_int64, C.POINTER(C.c_int64)]
    lib.pdhg_dev_finish.argtypes = [vp, _f64p, _f64p, _f64p, _f64p]
    lib.sankhya_pdhg_csr.argtypes = [vp, C.c_int, _i32p, _i32p, _f64p]
    lib.sankhya_pdhg_quality.argtypes = [vp, _f64p, _f64p, _f64p, C.c_double, _f64p, _f64p]
    lib.sankhya_pdhg_quality_threaded.argtypes = [vp, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                                                  _f64p, _f64p]
    lib.sankhya_pdhg_batch_open.argtypes = [C.POINTER(vp), C.c_int32, C.c_int, C.c_int, C.POINTER(vp)]
    lib.pdhg_batch_scaling.argtypes = [vp] + [_f64p] * 8
    lib.pdhg_batch_destroy.argtypes = [vp]
    lib.sankhya_pdhg_power_start.argtypes = [vp, C.c_int64, _f64p]
    lib.pdhg_dev_precondition.argtypes = [vp, C.POINTER(C.c_double)]
    lib.pdhg_dev_scaling.argtypes = [vp] + [_f64p] * 8
    lib.pdhg_dev_time_kernels.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.pdhg_dev_bytes.restype = C.c_int64
    lib.pdhg_dev_bytes.argtypes = [vp]
    lib.pdhg_dev_inject_fault.argtypes = [vp, C.c_int]
    _lib = lib
    return lib


def device_description(device: int = 0):
    """(available, text) for a gfx950 device, as the engine's own probe sees it."""
    lib = load_library()
    buf = C.create_string_buffer(256)
    rc = lib.pdhg_dev_probe(device, buf, 256)
    return rc == 0, buf.value.decode()


class Model:
    """An LP:  min/max offset + c.x  s.t.  row_lower <= A x <= row_upper,  col_lower <= x <= col_upper.

    A is CSC (colptr, rowidx ascending per column, val) like the reference's SparseMatrix.
    """

    def __init__(self, handle: int, owned: bool = True):
        self._lib = load_library()
        if not handle:
            raise ValueError("model construction failed")
        self._h = C.c_void_p(handle)
        self._owned = owned

    @classmethod
    def from_csc(cls, m, n, colptr, rowidx, val, cost, col_lower, col_upper, row_lower,
                 row_upper, objective_offset=0.0, maximize=False) -> "Model":
        lib = load_library()
        a = lambda v, t=np.float64: np.ascontiguousarray(v, dtype=t)
        colptr, rowidx = a(colptr, np.int32), a(rowidx, np.int32)
        h = lib.sankhya_model_new(int(m), int(n), int(colptr[-1]) if n else 0, colptr, rowidx, a(val),
                                  a(cost), a(col_lower), a(col_upper), a(row_lower), a(row_upper),
                                  float(objective_offset), 1 if maximize else 0)
        return cls(h)

    @classmethod
    def from_dense(cls, rows, row_lower, row_upper, cost, col_lower=None, col_upper=None,
                   objective_offset=0.0, maximize=False) -> "Model":
        A = np.asarray(rows, dtype=np.float64).reshape(len(row_lower), len(cost))
        m, n = A.shape
        colptr, rowidx, val = [0], [], []
        for j in range(n):
            for i in range(m):
                if A[i, j] != 0.0:
                    rowidx.append(i)
                    val.append(A[i, j])
            colptr.append(len(rowidx))
        cl = np.zeros(n) if col_lower is None else col_lower
        cu = np.full(n, np.inf) if col_upper is None else col_upper
        return cls.from_csc(m, n, colptr, rowidx, val, cost, cl, cu, row_lower, row_upper,
                            objective_offset, maximize)

    @classmethod
    def read_mps(cls, path: str) -> "Model":
        """An MPS file (free or fixed) read by the library's reader (sankhya_model_read)."""
        lib = load_library()
        h = lib.sankhya_model_create()
        model = cls(h)
        if lib.sankhya_model_read(model._h, os.fspath(path).encode()) != 0:
            raise ValueError(lib.sankhya_last_error().decode())
        return model

    def names(self) -> Dict[str, object]:
        """The model's name, source path and column and row names (empty strings when unset)."""
        m, n, _ = self.dims
        get = lambda kind, i=0: (self._lib.sankhya_model_name_of(self._h, kind, i) or b"").decode()
        return dict(name=get(0), source_path=get(1), col_names=[get(2, j) for j in range(n)],
                    row_names=[get(3, i) for i in range(m)])

    def integer_columns(self) -> np.ndarray:
        """Indices of the columns marked integer (the PDHG engine solves the relaxation)."""
        n = self.dims[1]
        return np.array([j for j in range(n) if self._lib.sankhya_model_col_is_integer(self._h, j) == 1],
                        dtype=np.int64)

    @classmethod
    def generate(cls, kind: str, m: int, n: int, per_col: int, seed: int, param: float = 1.5) -> "Model":
        kinds = {"uniform": 0, "zipf": 1, "blend": 2}
        lib = load_library()
        return cls(lib.sankhya_model_generate(kinds[kind], m, n, per_col, param, seed))

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "_h", None):
            self._lib.sankhya_model_free(self._h)
            self._h = None

    @property
    def dims(self):
        m, n, nnz = C.c_int32(), C.c_int32(), C.c_int64()
        self._lib.sankhya_model_dims(self._h, C.byref(m), C.byref(n), C.byref(nnz))
        return m.value, n.value, nnz.value

    def set_names(self, name=None, source_path=None, col_names=None, row_names=None) -> None:
        """Model name, source path and row/column names, as a solution file writes them."""
        m, n, _ = self.dims

        def names(seq, count):
            if seq is None:
                return None
            if len(seq) != count:
                raise ValueError(f"expected {count} names, got {len(seq)}")
            return (C.c_char_p * max(count, 1))(*[None if v is None else v.encode() for v in seq])

        self._lib.sankhya_model_set_names(self._h, None if name is None else name.encode(),
                                          None if source_path is None else source_path.encode(),
                                          names(col_names, n), names(row_names, m))

    def arrays(self) -> Dict[str, np.ndarray]:
        m, n, nnz = self.dims
        d = dict(colptr=np.zeros(n + 1, np.int32), rowidx=np.zeros(max(nnz, 1), np.int32),
                 val=np.zeros(max(nnz, 1)), cost=np.zeros(max(n, 1)), col_lower=np.zeros(max(n, 1)),
                 col_upper=np.zeros(max(n, 1)), row_lower=np.zeros(max(m, 1)),
                 row_upper=np.zeros(max(m, 1)))
        off, mx = C.c_double(), C.c_int32()
        self._lib.sankhya_model_export(self._h, d["colptr"], d["rowidx"], d["val"], d["cost"],
                                       d["col_lower"], d["col_upper"], d["row_lower"],
                                       d["row_upper"], C.byref(off), C.byref(mx))
        for k, ln in (("rowidx", nnz), ("val", nnz), ("cost", n), ("col_lower", n),
                      ("col_upper", n), ("row_lower", m), ("row_upper", m)):
            d[k] = d[k][:ln].copy()
        d.update(m=m, n=n, nnz=nnz, objective_offset=off.value, maximize=bool(mx.value))
        return d


@dataclass
class Solution:
    status: str
    objective: float
    dual_bound: float
    iterations: int
    col_value: np.ndarray
    row_dual: np.ndarray
    col_dual: np.ndarray
    row_activity: np.ndarray
    algorithm: str
    message: str
    solve_seconds: float
    primal_infeasibility: float
    dual_infeasibility: float
    complementarity_violation: float
    relative_gap: float
    stats: Dict[str, float] = field(default_factory=dict)

    @property
    def relative_kkt(self) -> float:
        """The relative KKT error max(relative primal, dual, gap) at the reported point."""
        return max(self.stats.get("final_primal", math.inf), self.stats.get("final_dual", math.inf),
                   self.stats.get("final_gap", math.inf))

    def reached(self, tolerance: float) -> bool:
        """True when the status claims the request (optimal, feasible) or, for a run under
        pdhg_stop_rule=relative_kkt, when the relative KKT error is at most `tolerance`. The
        second case claims no feasibility: its absolute primal residual may exceed 1e-7."""
        return self.status in ("optimal", "feasible") or self.relative_kkt <= tolerance


def _options_text(options: Dict[str, object]) -> bytes:
    parts = []
    for k, v in options.items():
        if isinstance(v, bool):
            v = "true" if v else "false"
        parts.append(f"{k}={v}")
    return ";".join(parts).encode()


def solve(model: Model, **options) -> Solution:
    """Solve with the MI355X PDHG engine. Options are the registry's names (pdhg_tolerance,
    iteration_limit, pdhg_stop_at_request, pdhg_precision, ...)."""
    lib = load_library()
    m, n, _ = model.dims
    x, y, d, act = np.zeros(max(n, 1)), np.zeros(max(m, 1)), np.zeros(max(n, 1)), np.zeros(max(m, 1))
    res = Result()
    rc = lib.sankhya_pdhg_solve(model._h, _options_text(options), x.ctypes.data, y.ctypes.data,
                                d.ctypes.data, act.ctypes.data, C.byref(res))
    if rc != 0:
        raise ValueError(res.message.decode())
    stats = {k: getattr(res, k) for k in ("attempts", "restarts", "ticks", "device_bytes", "fp32_iterations",
                                          "setup_seconds", "loop_seconds", "spectral_norm",
                                          "final_primal", "final_dual", "final_gap", "final_absolute_primal")}
    return Solution(STATUS_NAMES[res.status], res.objective, res.dual_bound, res.iterations,
                    x[:n], y[:m], d[:n], act[:m], res.algorithm.decode(), res.message.decode(),
                    res.solve_seconds, res.primal_infeasibility, res.dual_infeasibility,
                    res.complementarity_violation, res.relative_gap, stats)


def write_solution(model: Model, solution: Solution, path, **options) -> None:
    """Write the solution file (the reference's .sol format, `certificate none`) for a
    solution of `model`. Raises OSError when the file cannot be written."""
    lib = load_library()
    m, n, _ = model.dims
    res = Result()
    res.status = STATUS_NAMES.index(solution.status)
    res.objective, res.dual_bound = solution.objective, solution.dual_bound
    res.iterations, res.solve_seconds = solution.iterations, solution.solve_seconds
    res.primal_infeasibility = solution.primal_infeasibility
    res.dual_infeasibility = solution.dual_infeasibility
    res.algorithm = solution.algorithm.encode()[:31]
    res.message = solution.message.encode()[:1023]
    vec = lambda v, k: np.ascontiguousarray(v, dtype=np.float64) if v is not None and k else None
    x, d = vec(solution.col_value, n), vec(solution.col_dual, n)
    y, act = vec(solution.row_dual, m), vec(solution.row_activity, m)
    ptr = lambda v: None if v is None else v.ctypes.data
    err = C.create_string_buffer(1024)
    rc = lib.sankhya_write_solution(model._h, _options_text(options), str(path).encode(),
                                    C.byref(res), ptr(x), ptr(y), ptr(d), ptr(act), err, 1024)
    if rc != 0:
        raise OSError(err.value.decode())


def _solution_of(res, x, y, d, act, stats_extra=None) -> Solution:
    stats = {k: getattr(res, k) for k in ("attempts", "restarts", "ticks", "device_bytes", "fp32_iterations",
                                          "setup_seconds", "loop_seconds", "spectral_norm",
                                          "final_primal", "final_dual", "final_gap", "final_absolute_primal", "schedule")}
    return Solution(STATUS_NAMES[res.status], res.objective, res.dual_bound, res.iterations,
                    x, y, d, act, res.algorithm.decode(), res.message.decode(),
                    res.solve_seconds, res.primal_infeasibility, res.dual_infeasibility,
                    res.complementarity_violation, res.relative_gap, stats)


class BatchSession:
    """A solved batch kept on the device (sankhya_batch_solve) so that vector-Jacobian products
    of the LP solution maps can be taken afterwards (sankhya_batch_adjoint)."""

    def __init__(self, models, **options):
        self._lib = load_library()
        self.models = list(models)
        nlp = len(self.models)
        self.dims = [md.dims for md in self.models]
        sn = sum(d[1] for d in self.dims)
        sm = sum(d[0] for d in self.dims)
        x, d, y, act = np.zeros(max(sn, 1)), np.zeros(max(sn, 1)), np.zeros(max(sm, 1)), np.zeros(max(sm, 1))
        res = (Result * max(nlp, 1))()
        handles = (C.c_void_p * max(nlp, 1))(*[md._h.value for md in self.models])
        h = self._lib.sankhya_batch_solve(handles, nlp, _options_text(options), x.ctypes.data,
                                          y.ctypes.data, d.ctypes.data, act.ctypes.data, res)
        if not h:
            raise ValueError(res[0].message.decode())
        self._h = C.c_void_p(h)
        self.solutions, on, om = [], 0, 0
        for k, (m, n, _) in enumerate(self.dims):
            self.solutions.append(_solution_of(res[k], x[on:on + n].copy(), y[om:om + m].copy(),
                                               d[on:on + n].copy(), act[om:om + m].copy()))
            on += n
            om += m

    def adjoint(self, grad_col_value, grad_row_dual, tol=1e-14, max_iterations=50000):
        """dL/dc and dL/d(active row bound) per LP for upstream dL/dx and dL/d(row_dual) (lists
        of arrays, one per LP, or None for zero). Returns (grad_c, grad_b, iterations, residuals)."""
        def cat(parts, idx):
            out = []
            for k, dm in enumerate(self.dims):
                v = None if parts is None else parts[k]
                out.append(np.zeros(dm[idx]) if v is None else np.asarray(v, dtype=np.float64).reshape(dm[idx]))
            return np.ascontiguousarray(np.concatenate(out) if out else np.zeros(1))
        gx, gy = cat(grad_col_value, 1), cat(grad_row_dual, 0)
        gc, gb = np.zeros(max(gx.size, 1)), np.zeros(max(gy.size, 1))
        its = np.zeros(2 * max(len(self.dims), 1), dtype=np.int32)
        res = np.zeros(2 * max(len(self.dims), 1))
        rc = self._lib.sankhya_batch_adjoint(self._h, gx, gy, float(tol), int(max_iterations), gc, gb, its, res)
        if rc != 0:
            raise RuntimeError(f"adjoint failed with code {rc}")
        out_c, out_b, on, om = [], [], 0, 0
        for m, n, _ in self.dims:
            out_c.append(gc[on:on + n].copy())
            out_b.append(gb[om:om + m].copy())
            on += n
            om += m
        return out_c, out_b, its.reshape(-1, 2)[: len(self.dims)], res.reshape(-1, 2)[: len(self.dims)]

    def adjoint_masks(self):
        """The active set of the last adjoint call: (free-column masks, active-row masks), one
        pair of 0/1 arrays per LP."""
        sn, sm = sum(dm[1] for dm in self.dims), sum(dm[0] for dm in self.dims)
        mc, mr = np.zeros(max(sn, 1)), np.zeros(max(sm, 1))
        rc = self._lib.sankhya_batch_adjoint_masks(self._h, mc, mr)
        if rc != 0:
            raise RuntimeError(f"adjoint masks failed with code {rc}")
        out_c, out_r, on, om = [], [], 0, 0
        for m, n, _ in self.dims:
            out_c.append(mc[on:on + n].copy())
            out_r.append(mr[om:om + m].copy())
            on += n
            om += m
        return out_c, out_r

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sankhya_batch_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def solve_batch(models, **options):
    """Solve independent LPs together with the batched engine (sankhya_pdhg_solve_batch).
    Options as for solve(), plus pdhg_batch_tune and pdhg_batch_tune_attempts. Returns one
    Solution per model, in order."""
    lib = load_library()
    models = list(models)
    nlp = len(models)
    dims = [md.dims for md in models]
    sn = sum(d[1] for d in dims)
    sm = sum(d[0] for d in dims)
    x, d, y, act = np.zeros(max(sn, 1)), np.zeros(max(sn, 1)), np.zeros(max(sm, 1)), np.zeros(max(sm, 1))
    res = (Result * max(nlp, 1))()
    handles = (C.c_void_p * max(nlp, 1))(*[md._h.value for md in models])
    rc = lib.sankhya_pdhg_solve_batch(handles, nlp, _options_text(options), x.ctypes.data, y.ctypes.data,
                                      d.ctypes.data, act.ctypes.data, res)
    if rc != 0:
        raise ValueError(res[0].message.decode())
    out, on, om = [], 0, 0
    for k, (m, n, _) in enumerate(dims):
        out.append(_solution_of(res[k], x[on:on + n].copy(), y[om:om + m].copy(), d[on:on + n].copy(),
                                act[om:om + m].copy()))
        on += n
        om += m
    return out


def power_start(model: Model, random_seed: int = 0) -> np.ndarray:
    """The start vector of the engine's power iteration for this model (sankhya_pdhg_power_start)."""
    lib = load_library()
    n = model.dims[1]
    v = np.zeros(max(n, 1))
    if lib.sankhya_pdhg_power_start(model._h, int(random_seed), v) != 0:
        raise RuntimeError("sankhya_pdhg_power_start failed")
    return v[:n]


def host_csr(model: Model, threads: int = 1) -> Dict[str, np.ndarray]:
    """The CSR of the scaled matrix the engine uploads (sankhya_pdhg_csr), built on `threads`
    host threads."""
    lib = load_library()
    m, n, nnz = model.dims
    out = dict(rowptr=np.zeros(m + 1, np.int32), colidx=np.zeros(max(nnz, 1), np.int32),
               values=np.zeros(max(nnz, 1)))
    rc = lib.sankhya_pdhg_csr(model._h, int(threads), out["rowptr"], out["colidx"], out["values"])
    if rc != 0:
        raise RuntimeError(f"sankhya_pdhg_csr failed ({rc})")
    out["colidx"], out["values"] = out["colidx"][:nnz], out["values"][:nnz]
    return out


def quality(model: Model, col_value, row_dual, col_dual, dual_bound: float, threads: int = 0):
    """The engine's Solution::recompute_quality on the host (sankhya_pdhg_quality): the row
    activity and [objective, primal_infeasibility, primal_infeasibility_scaled,
    dual_infeasibility, dual_infeasibility_scaled, complementarity_violation,
    integrality_violation, absolute_gap, relative_gap]. With threads > 0, the threaded form a
    device solve's report uses (sankhya_pdhg_quality_threaded)."""
    lib = load_library()
    m, n, _ = model.dims
    act, out = np.zeros(max(m, 1)), np.zeros(9)
    f = lambda v, k: np.ascontiguousarray(v if k else np.zeros(1), np.float64)
    if threads > 0:
        rc = lib.sankhya_pdhg_quality_threaded(model._h, f(col_value, n), f(row_dual, m),
                                               f(col_dual, n), float(dual_bound), int(threads),
                                               act, out)
    else:
        rc = lib.sankhya_pdhg_quality(model._h, f(col_value, n), f(row_dual, m), f(col_dual, n),
                                      float(dual_bound), act, out)
    if rc != 0:
        raise RuntimeError("sankhya_pdhg_quality failed")
    return act[:m], out


def host_scaling(model: Model, ruiz: int = 10, threads: int = 1) -> Dict[str, np.ndarray]:
    """The engine's host preconditioner (Dr, Dc, scaled CSC values, scaled vectors), on
    `threads` host threads."""
    lib = load_library()
    m, n, nnz = model.dims
    out = dict(row=np.zeros(max(m, 1)), column=np.zeros(max(n, 1)), val=np.zeros(max(nnz, 1)),
               cost=np.zeros(max(n, 1)), col_lower=np.zeros(max(n, 1)), col_upper=np.zeros(max(n, 1)),
               row_lower=np.zeros(max(m, 1)), row_upper=np.zeros(max(m, 1)))
    lib.sankhya_pdhg_scaling(model._h, ruiz, int(threads), out["row"], out["column"], out["val"], out["cost"],
                             out["col_lower"], out["col_upper"], out["row_lower"], out["row_upper"])
    for k, ln in (("row", m), ("column", n), ("val", nnz), ("cost", n), ("col_lower", n),
                  ("col_upper", n), ("row_lower", m), ("row_upper", m)):
        out[k] = out[k][:ln]
    return out


def batch_device_scaling(models, precision: str = "fp64", device: int = 0) -> Dict[str, np.ndarray]:
    """The batched engine's device preconditioning of `models` (pdhg_batch_upload_unscaled),
    read back with pdhg_batch_scaling: row (Dr), column (Dc), val (CSC order), cost, col_lower,
    col_upper, row_lower, row_upper, each concatenated LP after LP."""
    lib = load_library()
    arr = (C.c_void_p * len(models))(*[m._h for m in models])
    h = C.c_void_p()
    rc = lib.sankhya_pdhg_batch_open(arr, len(models), device, PRECISION[precision], C.byref(h))
    if rc != 0:
        raise RuntimeError(f"batch open failed with code {rc}")
    sm = sum(m.dims[0] for m in models)
    sn = sum(m.dims[1] for m in models)
    sz = sum(m.dims[2] for m in models)
    out = {k: np.zeros(max(c, 1)) for k, c in (("row", sm), ("column", sn), ("val", sz), ("cost", sn),
                                              ("col_lower", sn), ("col_upper", sn), ("row_lower", sm),
                                              ("row_upper", sm))}
    try:
        rc = lib.pdhg_batch_scaling(h, out["row"], out["column"], out["val"], out["cost"],
                                    out["col_lower"], out["col_upper"], out["row_lower"], out["row_upper"])
        if rc != 0:
            raise RuntimeError(f"pdhg_batch_scaling failed ({rc})")
    finally:
        lib.pdhg_batch_destroy(h)
    cut = {"row": sm, "column": sn, "val": sz, "cost": sn, "col_lower": sn, "col_upper": sn,
           "row_lower": sm, "row_upper": sm}
    return {k: v[: cut[k]] for k, v in out.items()}


class Device:
    """A device handle holding one model as the engine uploads it; exposes the pdhg_dev_* hooks."""

    def __init__(self, model: Model, precision: str = "fp64", device: int = 0):
        self._lib = load_library()
        self.model = model
        self.m, self.n, self.nnz = model.dims
        h = C.c_void_p()
        rc = self._lib.sankhya_pdhg_device_open(model._h, device, PRECISION[precision], C.byref(h))
        if rc != 0:
            raise RuntimeError(f"device open failed with code {rc}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pdhg_dev_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self._lib.pdhg_dev_last_error(self._h).decode()}")

    def spmv(self, x: np.ndarray, transpose: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(max(self.n if transpose else self.m, 1))
        self._check(self._lib.pdhg_dev_spmv(self._h, 1 if transpose else 0, x, out), "spmv")
        return out[: (self.n if transpose else self.m)]

    def spectral_norm(self, v0: np.ndarray, iterations: int = 30) -> float:
        r = C.c_double()
        self._check(self._lib.pdhg_dev_spectral_norm(self._h, np.ascontiguousarray(v0, np.float64),
                                                     iterations, C.byref(r)), "spectral_norm")
        return r.value

    def evaluate(self, x: np.ndarray, y: np.ndarray):
        act, red, res = np.zeros(max(self.m, 1)), np.zeros(max(self.n, 1)), np.zeros(9)
        self._check(self._lib.pdhg_dev_evaluate(self._h, np.ascontiguousarray(x, np.float64),
                                                np.ascontiguousarray(y, np.float64), act, red, res),
                    "evaluate")
        return act[: self.m], red[: self.n], res

    def start(self, tolerance=1e-4, iteration_limit=1000000, restart=True, stop_at_request=False,
              spectral_norm=1.0, chunk=0, stop_rule=0):
        p = DevParams(tolerance, iteration_limit, 1 if restart else 0, 1 if stop_at_request else 0,
                      spectral_norm, chunk, stop_rule)
        self._check(self._lib.pdhg_dev_start(self._h, C.byref(p)), "start")

    def promote(self, tolerance=1e-8, iteration_limit=1000000, restart=True,
                stop_at_request=False, spectral_norm=1.0, chunk=0, stop_rule=0):
        """Mixed-precision handle: continue the same run in FP64 (pdhg_dev_promote)."""
        p = DevParams(tolerance, iteration_limit, 1 if restart else 0, 1 if stop_at_request else 0,
                      spectral_norm, chunk, stop_rule)
        self._check(self._lib.pdhg_dev_promote(self._h, C.byref(p)), "promote")

    def run(self, chunks: int = 1) -> DevState:
        st = DevState()
        self._check(self._lib.pdhg_dev_run(self._h, chunks, C.byref(st)), "run")
        return st

    def ticks(self, first: int = 0, max_count: int = 1 << 16):
        buf = (Tick * max_count)()
        got = C.c_int64()
        self._check(self._lib.pdhg_dev_ticks(self._h, first, buf, max_count, C.byref(got)), "ticks")
        return [buf[k] for k in range(got.value)]

    def finish(self):
        x, y, red, res = np.zeros(max(self.n, 1)), np.zeros(max(self.m, 1)), np.zeros(max(self.n, 1)), np.zeros(9)
        self._check(self._lib.pdhg_dev_finish(self._h, x, y, red, res), "finish")
        return x[: self.n], y[: self.m], red[: self.n], res

    def precondition(self):
        """Restore the LP's own values kept in HBM and precondition again on the device
        (pdhg_dev_precondition): the first step of a new solve of the resident LP."""
        ext = (C.c_double * 2)()
        self._check(self._lib.pdhg_dev_precondition(self._h, ext), "precondition")
        return ext[0], ext[1]

    def solve(self, tolerance: float = 1e-4, stop_rule: int = 0, stop_at_request: bool = False,
              iteration_limit: int = 1000000, random_seed: int = 0, chunk: int = 0):
        """One whole FP64 solve of the LP this handle holds in HBM, with nothing uploaded from
        the host: precondition again from the kept values, the power iteration from the engine's
        start vector, start, graph chunks until the device's stopping test or limit ends the run
        (the single-LP engine's own loop, src/pdhg_engine.cpp, without time limit or callback),
        then the final point. Returns (state, x, y, reduced, residuals); x and y are unscaled."""
        import time as _time
        self.precondition()
        norm = self.spectral_norm(power_start(self.model, random_seed), 30)
        self.start(tolerance=tolerance, iteration_limit=iteration_limit, restart=True,
                   stop_at_request=stop_at_request, spectral_norm=norm, chunk=chunk,
                   stop_rule=stop_rule)
        batch = 1
        while True:
            t0 = _time.perf_counter()
            st = self.run(batch)
            dt = _time.perf_counter() - t0
            if st.done or st.limit_hit or st.halted:
                break
            # Keep each wait near 20 ms, as the engine's loop does.
            if dt < 0.01 and batch < 256:
                batch *= 2
            elif dt > 0.04 and batch > 1:
                batch //= 2
        x, y, red, res = self.finish()
        return st, x, y, red, res

    def scaling(self) -> Dict[str, np.ndarray]:
        """The device's preconditioning (pdhg_dev_scaling): Dr, Dc, scaled CSC values, scaled
        cost and bounds, under the keys host_scaling() uses."""
        m, n, nnz = self.m, self.n, self.nnz
        out = dict(row=np.zeros(max(m, 1)), column=np.zeros(max(n, 1)), val=np.zeros(max(nnz, 1)),
                   cost=np.zeros(max(n, 1)), col_lower=np.zeros(max(n, 1)), col_upper=np.zeros(max(n, 1)),
                   row_lower=np.zeros(max(m, 1)), row_upper=np.zeros(max(m, 1)))
        self._check(self._lib.pdhg_dev_scaling(self._h, out["row"], out["column"], out["val"], out["cost"],
                                               out["col_lower"], out["col_upper"], out["row_lower"],
                                               out["row_upper"]), "scaling")
        for k, ln in (("row", m), ("column", n), ("val", nnz), ("cost", n), ("col_lower", n),
                      ("col_upper", n), ("row_lower", m), ("row_upper", m)):
            out[k] = out[k][:ln]
        return out

    def time_kernels(self, reps: int = 20):
        a, b = C.c_double(), C.c_double()
        self._check(self._lib.pdhg_dev_time_kernels(self._h, reps, C.byref(a), C.byref(b)), "time_kernels")
        return a.value, b.value

    @property
    def bytes(self) -> int:
        return int(self._lib.pdhg_dev_bytes(self._h))
