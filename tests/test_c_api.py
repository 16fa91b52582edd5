"""SANKHYA's C API (include/sankhya_c.h, the reference's include/sankhya/sankhya.h subset for
this path) and the VRAM check before allocation (src/core/solve.cpp:751-771,
src/gpu/gpu_memory.cpp:18-55).

CPU: the builder, option errors, validation, buffer-size contracts, the device-memory estimate
against a hand count and the fit decision with injected free sizes. GPU: solves through
sankhya_solve of the reference's own known answers (tests/unit/test_pdhg.cpp:95-146), the
progress callback and sankhya_model_interrupt ending a solve as interrupted
(src/core/stop_controller.hpp:58-71, src/pdhg/pdhg.cpp:208-218), the gate refusing an LP
before anything is allocated, and the estimate equal to the bytes a handle holds."""
import ctypes as C
import math
import os
import threading
import time

import numpy as np
import pytest

import common as cm
import sankhya_hip as sh

vp = C.c_void_p
PROGRESS = None


class Progress(C.Structure):
    _fields_ = [("phase", C.c_int), ("iterations", C.c_int64), ("nodes", C.c_int64),
                ("objective", C.c_double), ("best_bound", C.c_double), ("gap", C.c_double),
                ("elapsed_seconds", C.c_double), ("open_nodes", C.c_int64)]


CALLBACK = C.CFUNCTYPE(C.c_int, C.POINTER(Progress), vp)
STATUS = {0: "not_solved", 1: "optimal", 2: "feasible", 3: "infeasible", 4: "unbounded",
          10: "infeasible_or_unbounded", 5: "iteration_limit", 6: "time_limit", 7: "node_limit",
          8: "numerical_error", 9: "model_error", 11: "interrupted"}


def lib():
    L = sh.load_library()
    if getattr(L, "_c_api_declared", False):
        return L
    L.sankhya_last_error.restype = C.c_char_p
    L.sankhya_version.restype = C.c_char_p
    L.sankhya_infinity.restype = C.c_double
    L.sankhya_model_create.restype = vp
    L.sankhya_model_set_maximize.argtypes = [vp, C.c_int]
    L.sankhya_model_set_objective_offset.argtypes = [vp, C.c_double]
    L.sankhya_model_add_column.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_char_p,
                                           C.POINTER(C.c_int)]
    L.sankhya_model_add_row.argtypes = [vp, C.c_double, C.c_double, C.c_char_p, C.POINTER(C.c_int)]
    L.sankhya_model_set_coefficient.argtypes = [vp, C.c_int, C.c_int, C.c_double]
    L.sankhya_model_set_col_bounds.argtypes = [vp, C.c_int, C.c_double, C.c_double]
    L.sankhya_model_set_row_bounds.argtypes = [vp, C.c_int, C.c_double, C.c_double]
    L.sankhya_model_set_objective_coefficient.argtypes = [vp, C.c_int, C.c_double]
    for f in ("num_cols", "num_rows", "num_nonzeros"):
        getattr(L, "sankhya_model_" + f).argtypes = [vp]
    L.sankhya_model_validate.argtypes = [vp]
    L.sankhya_set_callback.argtypes = [vp, CALLBACK, vp]
    L.sankhya_model_interrupt.argtypes = [vp]
    L.sankhya_options_create.restype = vp
    L.sankhya_options_free.argtypes = [vp]
    L.sankhya_options_set_bool.argtypes = [vp, C.c_char_p, C.c_int]
    L.sankhya_options_set_int.argtypes = [vp, C.c_char_p, C.c_int64]
    L.sankhya_options_set_double.argtypes = [vp, C.c_char_p, C.c_double]
    L.sankhya_options_set_string.argtypes = [vp, C.c_char_p, C.c_char_p]
    L.sankhya_solve.argtypes = [vp, vp, C.POINTER(vp)]
    L.sankhya_solution_free.argtypes = [vp]
    L.sankhya_solution_status.argtypes = [vp]
    for f in ("message", "algorithm"):
        getattr(L, "sankhya_solution_" + f).argtypes = [vp]
        getattr(L, "sankhya_solution_" + f).restype = C.c_char_p
    for f in ("objective", "dual_bound", "relative_gap", "seconds", "primal_infeasibility",
              "dual_infeasibility"):
        getattr(L, "sankhya_solution_" + f).argtypes = [vp]
        getattr(L, "sankhya_solution_" + f).restype = C.c_double
    L.sankhya_solution_iterations.argtypes = [vp]
    L.sankhya_solution_iterations.restype = C.c_int64
    for f in ("col_values", "row_activities", "row_duals", "col_duals"):
        getattr(L, "sankhya_solution_" + f).argtypes = [vp, vp, C.c_int]
    L.sankhya_solution_claims_a_point.argtypes = [vp]
    L.sankhya_pdhg_memory_check.argtypes = [vp, C.c_int, C.c_int64, C.c_int64, C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int64)]
    L.pdhg_dev_estimate_bytes.restype = C.c_int64
    L.pdhg_dev_estimate_bytes.argtypes = [C.c_int32, C.c_int32, C.c_int64, vp, vp, vp, vp, C.c_int,
                                          C.c_int, C.c_int]
    L.pdhg_dev_vram_reserve.restype = C.c_int64
    L.pdhg_dev_vram_reserve.argtypes = [C.c_int64]
    L.pdhg_dev_bytes.restype = C.c_int64
    L.pdhg_dev_bytes.argtypes = [vp]
    L._c_api_declared = True
    return L


def build(L, rows, lower, upper, cost, col_lower=None, col_upper=None, maximize=False, offset=0.0):
    """The reference's make_lp (tests/unit/test_pdhg.cpp:66-89) through the builder calls."""
    A = np.asarray(rows, dtype=float)
    m, n = A.shape
    h = L.sankhya_model_create()
    assert h
    inf = L.sankhya_infinity()
    for j in range(n):
        lo = 0.0 if col_lower is None else col_lower[j]
        hi = inf if col_upper is None else col_upper[j]
        idx = C.c_int()
        assert L.sankhya_model_add_column(h, cost[j], lo, hi, 0, None, C.byref(idx)) == 0 and idx.value == j
    for i in range(m):
        assert L.sankhya_model_add_row(h, lower[i], upper[i], None, None) == 0
    for i in range(m):
        for j in range(n):
            if A[i, j] != 0.0:
                assert L.sankhya_model_set_coefficient(h, i, j, A[i, j]) == 0
    if maximize:
        L.sankhya_model_set_maximize(h, 1)
    if offset:
        L.sankhya_model_set_objective_offset(h, offset)
    return h


def test_builder_validation_and_option_errors():
    L = lib()
    h = build(L, [[1.0, 1.0]], [1.0], [math.inf], [1.0, 1.0])
    assert (L.sankhya_model_num_rows(h), L.sankhya_model_num_cols(h), L.sankhya_model_num_nonzeros(h)) == (1, 2, 2)
    # Replacing an entry keeps one; an exact zero removes it (sankhya.h:133-141).
    assert L.sankhya_model_set_coefficient(h, 0, 1, 3.0) == 0 and L.sankhya_model_num_nonzeros(h) == 2
    assert L.sankhya_model_set_coefficient(h, 0, 1, 0.0) == 0 and L.sankhya_model_num_nonzeros(h) == 1
    assert L.sankhya_model_set_coefficient(h, 5, 0, 1.0) == 1  # SANKHYA_ERROR_ARGUMENT
    assert b"outside the model" in L.sankhya_last_error()
    assert L.sankhya_model_validate(h) == 0
    L.sankhya_model_set_col_bounds(h, 0, 2.0, 1.0)  # lower above upper: a model error
    assert L.sankhya_model_validate(h) == 3 and L.sankhya_last_error()
    o = L.sankhya_options_create()
    assert L.sankhya_options_set_double(o, b"pdhg_tolerance", 1e-6) == 0
    assert L.sankhya_options_set_double(o, b"pdhg_tolerance", 2.0) == 4  # out of range
    assert L.sankhya_options_set_int(o, b"no_such_option", 1) == 4
    assert b"unknown option" in L.sankhya_last_error()
    assert L.sankhya_options_set_string(o, b"pdhg_precision", b"fp16") == 4
    assert L.sankhya_options_set_bool(o, b"gpu", 1) == 0
    L.sankhya_options_free(o)
    sh.load_library().sankhya_model_free(h)


def test_solution_getters_refuse_a_wrong_size():
    L = lib()
    assert L.sankhya_solution_col_values(None, None, 0) == 1
    assert L.sankhya_solution_claims_a_point(None) == 0 and L.sankhya_solution_status(None) == 0


def _bytes(count, elem):
    return (count + 2) * elem  # every device array carries two spare elements


def test_device_memory_estimate_against_a_hand_count():
    """m = 2, n = 3, 4 nonzeros: one stream item per matrix, no long rows. The count follows the
    arrays a handle allocates, listed by hand (FP64, uploaded unscaled, values not kept)."""
    L = lib()
    d = cm.dense_lp([[1.0, 2.0, 0.0], [0.0, 3.0, 4.0]], [1.0, 1.0], [1.0, 1.0], [1.0, 1.0, 1.0])
    mdl = cm.model_of(d)
    csr = sh.host_csr(mdl)
    m, n, nnz = 2, 3, 4
    ctl, tick = 304, 48  # sizeof(Ctl), sizeof(pdhg_tick)
    # row pointers, indices, values, one item, no long rows / tail groups / segment sums /
    # staged long-row inputs
    plan = lambda rows: (_bytes(rows + 1, 4) + _bytes(nnz, 4) + _bytes(nnz, 8) + _bytes(1, 16)
                         + _bytes(0, 16) + _bytes(0, 16) + _bytes(0, 8) + _bytes(0, 8))
    want = (plan(m) + plan(n) + _bytes(2 * n, 8) + _bytes(2 * m, 8) + 6 * _bytes(n, 8) + 6 * _bytes(m, 8)
            # partials of k_primal and k_dual padded to one round of the tail's loads (4 x 1024),
            # the long-column and long-row sums, the tick partials (9 per item, same padding)
            + 7 * _bytes(n, 8) + 5 * _bytes(m, 8) + _bytes(4096, 8) + _bytes(1, 8) + _bytes(8192, 8)
            + _bytes(1, 8) + _bytes(9 * 4096, 8) + _bytes(9 * 4096, 8) + _bytes(1, ctl)
            + _bytes(1 << 16, tick)
            + _bytes(2, 8)
            + _bytes(2 * 1024 + 2, 8))
    got = L.pdhg_dev_estimate_bytes(m, n, nnz, csr["rowptr"].ctypes.data, csr["colidx"].ctypes.data,
                                    d["colptr"].ctypes.data, d["rowidx"].ctypes.data, 0, 1, 0)
    assert got == want
    kept = L.pdhg_dev_estimate_bytes(m, n, nnz, csr["rowptr"].ctypes.data, csr["colidx"].ctypes.data,
                                     d["colptr"].ctypes.data, d["rowidx"].ctypes.data, 0, 1, 1)
    assert kept == want + 2 * _bytes(nnz, 8)
    mixed = L.pdhg_dev_estimate_bytes(m, n, nnz, csr["rowptr"].ctypes.data, csr["colidx"].ctypes.data,
                                      d["colptr"].ctypes.data, d["rowidx"].ctypes.data, 1, 1, 0)
    # FP32 values beside the FP64 ones, FP32 iterate pairs, and the FP64 pairs of the handoff.
    assert mixed == want + 2 * _bytes(nnz, 4) - _bytes(2 * n, 8) - _bytes(2 * m, 8) \
        + _bytes(2 * n, 4) + _bytes(2 * m, 4) + _bytes(2 * n, 8) + _bytes(2 * m, 8)
    bad = np.array([0, 1, 9], dtype=np.int32)  # row pointer that does not end at nnz
    assert L.pdhg_dev_estimate_bytes(m, n, nnz, bad.ctypes.data, csr["colidx"].ctypes.data,
                                     d["colptr"].ctypes.data, d["rowidx"].ctypes.data, 0, 1, 0) == -1


def test_estimate_grows_with_the_lp_and_the_fit_decision():
    L = lib()
    sizes = []
    for m, n in ((200, 600), (400, 1200), (800, 2400)):
        mdl = sh.Model.generate("uniform", m, n, 5, 1)
        req, res = C.c_int64(), C.c_int64()
        assert L.sankhya_pdhg_memory_check(mdl._h, 0, 1 << 40, 1 << 40, C.byref(req), C.byref(res)) == 1
        sizes.append(req.value)
    assert sizes[0] < sizes[1] < sizes[2]
    # The reserve is max(256 MiB, total / 10) (src/gpu/gpu_memory.cpp:50-55).
    assert L.pdhg_dev_vram_reserve(0) == 256 << 20
    assert L.pdhg_dev_vram_reserve(288 << 30) == (288 << 30) // 10
    mdl = sh.Model.generate("zipf", 2000, 5000, 10, 3)
    req, res = C.c_int64(), C.c_int64()
    total = 8 << 30
    assert L.sankhya_pdhg_memory_check(mdl._h, 0, total, total, C.byref(req), C.byref(res)) == 1
    reserve = L.pdhg_dev_vram_reserve(total)
    assert res.value == reserve
    # Exactly at the boundary it fits; one byte short it does not.
    assert L.sankhya_pdhg_memory_check(mdl._h, 0, req.value + reserve, total, None, None) == 1
    assert L.sankhya_pdhg_memory_check(mdl._h, 0, req.value + reserve - 1, total, None, None) == 0


# ---- GPU -------------------------------------------------------------------------------------

def _solve(L, h, **opts):
    o = L.sankhya_options_create()
    for k, v in dict(dict(log_to_console=False), **opts).items():
        if isinstance(v, bool):
            assert L.sankhya_options_set_bool(o, k.encode(), int(v)) == 0
        elif isinstance(v, int):
            assert L.sankhya_options_set_int(o, k.encode(), v) == 0
        elif isinstance(v, float):
            assert L.sankhya_options_set_double(o, k.encode(), v) == 0
        else:
            assert L.sankhya_options_set_string(o, k.encode(), v.encode()) == 0
    sol = vp()
    assert L.sankhya_solve(h, o, C.byref(sol)) == 0
    L.sankhya_options_free(o)
    return sol


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    # tests/unit/test_pdhg.cpp:95-146, built with make_lp's defaults (x >= 0, no upper bound).
    ("MinimisesASum", [[1.0, 1.0]], [2.0], [math.inf], [1.0, 1.0], None, False, 0.0, 2.0),
    ("HonoursAnEqualityRow", [[1.0, 1.0]], [4.0], [4.0], [1.0, 2.0], None, False, 0.0, 4.0),
    ("HonoursARangeRow", [[1.0]], [2.0], [6.0], [1.0], None, False, 0.0, 2.0),
    ("HonoursAColumnUpperBound", [[1.0]], [-math.inf], [100.0], [-1.0], [7.0], False, 0.0, -7.0),
    ("ReportsTheObjectiveInTheOriginalSense", [[1.0, 0.0], [0.0, 2.0], [3.0, 2.0]],
     [-math.inf] * 3, [4.0, 12.0, 18.0], [3.0, 5.0], None, True, 0.0, 36.0),
    ("CarriesTheObjectiveOffset", [[1.0]], [3.0], [math.inf], [1.0], None, False, 10.0, 13.0),
], ids=lambda c: c[0])
def test_c_api_solves_the_reference_known_answers(case):
    cm.gpu_or_skip()
    L = lib()
    name, rows, lo, hi, cost, cu, maximize, offset, want = case
    h = build(L, rows, lo, hi, cost, col_upper=cu, maximize=maximize, offset=offset)
    sol = _solve(L, h, pdhg_tolerance=1e-8, iteration_limit=200000, pdhg_polish=False)
    status = STATUS[L.sankhya_solution_status(sol)]
    obj = L.sankhya_solution_objective(sol)
    print(f"{name}: {status} {obj:.12g} ({L.sankhya_solution_algorithm(sol).decode()})", flush=True)
    assert status == "optimal" and abs(obj - want) <= 1e-5 * max(1.0, abs(want))
    n, m = len(cost), len(rows)
    x = np.zeros(n)
    assert L.sankhya_solution_col_values(sol, x.ctypes.data, n) == 0
    assert L.sankhya_solution_col_values(sol, x.ctypes.data, n + 1) == 1  # wrong size refused
    y, act, dj = np.zeros(m), np.zeros(m), np.zeros(n)
    for f, buf, k in (("row_duals", y, m), ("row_activities", act, m), ("col_duals", dj, n)):
        assert getattr(L, "sankhya_solution_" + f)(sol, buf.ctypes.data, k) == 0
    assert np.allclose(act, np.asarray(rows) @ x, atol=1e-9)
    assert abs(offset + np.dot(cost, x) - obj) <= 1e-9 * max(1.0, abs(obj))
    assert L.sankhya_solution_claims_a_point(sol) == 1
    L.sankhya_solution_free(sol)
    sh.load_library().sankhya_model_free(h)


def _long_lp(L):
    """The 1,000 x 3,000 instance at a tolerance it takes many chunks to reach."""
    mdl, _ = cm.c0(1)
    return mdl


@pytest.mark.gpu
def test_progress_callback_returning_nonzero_interrupts_the_solve():
    cm.gpu_or_skip()
    L = lib()
    mdl = _long_lp(L)
    calls = []

    def cb(p, _user):
        calls.append((p.contents.iterations, p.contents.elapsed_seconds, p.contents.phase))
        return 1 if p.contents.iterations > 0 else 0

    fn = CALLBACK(cb)
    assert L.sankhya_set_callback(mdl._h, fn, None) == 0
    sol = _solve(L, mdl._h, pdhg_tolerance=1e-12, iteration_limit=1000000, pdhg_polish=False)
    assert L.sankhya_set_callback(mdl._h, CALLBACK(0), None) == 0
    status = STATUS[L.sankhya_solution_status(sol)]
    its = L.sankhya_solution_iterations(sol)
    print(f"callback calls {len(calls)}, first {calls[:2]}, status {status} after {its} iterations", flush=True)
    assert status == "interrupted" and L.sankhya_solution_claims_a_point(sol) == 1
    assert len(calls) >= 2 and calls[0][2] == 1  # SANKHYA_PHASE_LP
    assert 0 < its < 1000000
    L.sankhya_solution_free(sol)


@pytest.mark.gpu
def test_interrupt_from_another_thread_ends_the_solve():
    cm.gpu_or_skip()
    L = lib()
    mdl = _long_lp(L)
    out = {}

    def run():
        out["sol"] = _solve(L, mdl._h, pdhg_tolerance=1e-12, iteration_limit=1000000, pdhg_polish=False)

    t = threading.Thread(target=run)
    t0 = time.time()
    t.start()
    time.sleep(1.0)
    assert L.sankhya_model_interrupt(mdl._h) == 0
    t.join(timeout=120)
    assert not t.is_alive()
    sol = out["sol"]
    status = STATUS[L.sankhya_solution_status(sol)]
    its = L.sankhya_solution_iterations(sol)
    print(f"interrupted after {time.time()-t0:.2f}s: {status}, {its} iterations", flush=True)
    assert status == "interrupted" and 0 < its < 1000000
    L.sankhya_solution_free(sol)


@pytest.mark.gpu
def test_vram_gate_refuses_before_allocating(monkeypatch):
    cm.gpu_or_skip()
    L = lib()
    mdl, _ = cm.c0(1)
    monkeypatch.setenv("SANKHYA_HIP_FREE_BYTES", str(64 << 20))  # less than the reserve alone
    s = sh.solve(mdl, pdhg_tolerance=1e-6, iteration_limit=1000)
    print(s.status, s.message, flush=True)
    assert s.status == "numerical_error" and "nothing was allocated" in s.message
    monkeypatch.delenv("SANKHYA_HIP_FREE_BYTES")
    assert sh.solve(mdl, pdhg_tolerance=1e-6, iteration_limit=1000).status == "iteration_limit"


@pytest.mark.gpu
@pytest.mark.parametrize("name,make", [
    ("c0", lambda: cm.c0(1)),
    ("zipf-20000x50000", lambda: cm.generated("zipf", 20000, 50000, 10, 3)),
    ("transposed-zipf", lambda: cm.transposed_zipf(2000, 5000, 3)),
    ("afiro", lambda: (lambda d: (cm.model_of(d), d))(cm.netlib("afiro"))),
])
def test_estimate_equals_the_bytes_a_handle_holds(name, make):
    cm.gpu_or_skip()
    L = lib()
    mdl, d = make()
    m, n, nnz = mdl.dims
    csr = sh.host_csr(mdl)
    for precision in ("fp64", "mixed"):
        est = L.pdhg_dev_estimate_bytes(m, n, nnz, csr["rowptr"].ctypes.data, csr["colidx"].ctypes.data,
                                        d["colptr"].ctypes.data, d["rowidx"].ctypes.data,
                                        sh.PRECISION[precision], 1, 1)
        dev = sh.Device(mdl, precision=precision)
        if precision == "mixed":
            dev.start(tolerance=1e-4, spectral_norm=1.0, stop_rule=1, iteration_limit=10)
            dev.promote(tolerance=1e-4, spectral_norm=1.0, stop_rule=1, iteration_limit=20)
        held = L.pdhg_dev_bytes(dev._h)
        dev.close()
        print(f"{name} {precision}: estimate {est} bytes, handle holds {held}", flush=True)
        assert est == held
