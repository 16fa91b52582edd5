"""Pinning the CPU oracle (oracle/pdhg_oracle.c) before it is trusted as the checker.

Each oracle function is checked against something other than its own formula:
  - build_scaling, spectral_norm, evaluate: bit-exact against the REFERENCE's own functions,
    through committed golden vectors (tests/golden/ref_vectors.npz, written by
    tests/golden/make_golden.py from the reference sources compiled in place) and, where
    oracle/_ref/libref.so is built, directly on more instances;
  - std::mt19937 / uniform_real_distribution: the C++ standard's known answer and libstdc++;
  - SpMV: dense brute force and the adjoint identity x'(A'y) = (Ax)'y;
  - spectral norm: closed forms (diagonal matrix) and numpy's SVD;
  - scaling: the 1x1 closed form;
  - evaluate: a hand-computed optimal pair with zero residuals;
  - the whole engine (step rule, restarts, stopping test, report): the reference's own
    known-answer tests (tests/unit/test_pdhg.cpp) and the published Netlib optima
    (data/netlib/reference.json). Exact trajectories (iteration counts) are not pinned: the
    reference engine cannot be built here (its Logger needs nlohmann_json, which is absent).
A planted mistake (nine Ruiz rounds instead of ten) is shown to fail the golden check.
"""
import ctypes as C
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import common as cm
import oracle as orc
import sankhya_hip as sh

GOLDEN = os.path.join(cm.HERE, "golden", "ref_vectors.npz")
CASES = ["afiro", "blend", "uniform_100x300_s5", "zipf_300x800_s4"]


def golden_problem(g, name):
    m, n, nnz, mx = g[f"{name}/in_dims"]
    d = dict(m=int(m), n=int(n), objective_offset=float(g[f"{name}/in_offset"][0]), maximize=bool(mx))
    for k in ("colptr", "rowidx", "val", "cost", "col_lower", "col_upper", "row_lower", "row_upper"):
        d[k] = g[f"{name}/in_{k}"]
    return orc.Problem(d)


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def test_mt19937_known_answer():
    # C++ [rand.predef]: the 10000th invocation of a default-constructed mt19937 is 4123659995.
    g = orc.OrMt()
    orc.lib().or_mt19937_seed(C.byref(g), 5489)
    v = 0
    for _ in range(10000):
        v = orc.lib().or_mt19937_next(C.byref(g))
    assert v == 4123659995


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_uniform_real_matches_libstdcxx(tmp_path):
    src = tmp_path / "u.cpp"
    src.write_text("#include <cstdio>\n#include <random>\nint main(){std::mt19937 g(12345);"
                   "std::uniform_real_distribution<double> u(-1.0,1.0);"
                   "for(int i=0;i<64;++i)std::printf(\"%.17g\\n\",u(g));}\n")
    exe = tmp_path / "u"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    want = [float(t) for t in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    g = orc.OrMt()
    orc.lib().or_mt19937_seed(C.byref(g), 12345)
    got = [orc.lib().or_uniform_real(C.byref(g), -1.0, 1.0) for _ in range(64)]
    assert got == want


@pytest.mark.parametrize("name", CASES)
def test_scaling_matches_reference_golden(golden, name):
    P = golden_problem(golden, name)
    s = orc.build_scaling(P)
    for k, gk in (("row", "row"), ("column", "column"), ("val", "val"), ("cost", "cost"),
                  ("col_lower", "l"), ("col_upper", "u"), ("row_lower", "rl"), ("row_upper", "ru")):
        assert np.array_equal(s[k], golden[f"{name}/scaling_{gk}"]), k


@pytest.mark.parametrize("name", CASES)
def test_spectral_norm_matches_reference_golden(golden, name):
    P = golden_problem(golden, name)
    s = orc.build_scaling(P)
    assert orc.spectral_norm(P, s["val"], 30, 1) == golden[f"{name}/spectral_norm_seed1"][0]


@pytest.mark.parametrize("name", CASES)
def test_evaluate_matches_reference_golden(golden, name):
    P = golden_problem(golden, name)
    act, red, res = orc.evaluate(P, golden[f"{name}/eval_x"], golden[f"{name}/eval_y"])
    assert np.array_equal(act, golden[f"{name}/eval_activity"])
    assert np.array_equal(red, golden[f"{name}/eval_reduced"])
    assert np.array_equal(res, golden[f"{name}/eval_residuals"])


def test_planted_mistake_fails_the_golden_check(golden):
    P = golden_problem(golden, "afiro")
    wrong = orc.build_scaling(P, ruiz=9)
    assert not np.array_equal(wrong["val"], golden["afiro/scaling_val"])


def _ref_cases():
    yield "C0-s1", cm.c0(1)[1]
    yield "C0-s2", cm.c0(2)[1]
    yield "zipf-2000x5000", cm.generated("zipf", 2000, 5000, 10, 3)[1]
    for nm in ("afiro", "sc50a", "sc50b", "adlittle", "blend", "share2b", "sc105", "stocfor1", "israel"):
        yield nm, cm.netlib(nm)


# The reference library is opened inside the test body only: a skip condition runs at collection,
# which would load it in every run, GPU runs included, where it is never shipped.
@pytest.mark.skipif(not os.path.exists(orc.REF_SO), reason="oracle/_ref/libref.so not built (no reference tree)")
@pytest.mark.parametrize("name,d", list(_ref_cases()), ids=[c[0] for c in _ref_cases()])
def test_oracle_matches_reference_library(name, d):
    R = orc.ref()
    P = orc.Problem(d)
    s = orc.build_scaling(P)
    r = {k: np.zeros(max(sz, 1)) for k, sz in (("row", P.m), ("column", P.n), ("val", P.nnz), ("cost", P.n),
                                                 ("l", P.n), ("u", P.n), ("rl", P.m), ("ru", P.m))}
    R.ref_build_scaling(P.m, P.n, P.colptr, P.rowidx, P.val, P.cost, P.col_lower, P.col_upper,
                        P.row_lower, P.row_upper, P.min_cost, 10, r["row"], r["column"], r["val"],
                        r["cost"], r["l"], r["u"], r["rl"], r["ru"])
    assert np.array_equal(s["row"], r["row"][: P.m])
    assert np.array_equal(s["column"], r["column"][: P.n])
    assert np.array_equal(s["val"], r["val"][: P.nnz])
    for seed in (1, 7):
        assert orc.spectral_norm(P, s["val"], 30, seed) == R.ref_spectral_norm(P.m, P.n, P.colptr, P.rowidx, s["val"], 30, seed)
    rng = np.random.default_rng(3)
    x = np.ascontiguousarray(rng.uniform(-2, 5, P.n))
    y = np.ascontiguousarray(rng.standard_normal(P.m))
    a1, r1, res1 = orc.evaluate(P, x, y)
    a2, r2, res2 = np.zeros(max(P.m, 1)), np.zeros(max(P.n, 1)), np.zeros(9)
    R.ref_evaluate(P.m, P.n, P.colptr, P.rowidx, P.val, P.cost, P.col_lower, P.col_upper, P.row_lower,
                   P.row_upper, P.min_cost, P.bound_norm, P.cost_norm, x, y, a2, r2, res2)
    assert np.array_equal(a1, a2[: P.m]) and np.array_equal(r1, r2[: P.n]) and np.array_equal(res1, res2)


def test_spmv_brute_force_and_adjoint():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((7, 9)) * (rng.random((7, 9)) < 0.4)
    d = cm.dense_lp(A, [-np.inf] * 7, [np.inf] * 7, np.zeros(9))
    P = orc.Problem(d)
    x, y = rng.standard_normal(9), rng.standard_normal(7)
    assert np.allclose(orc.csc_multiply(P, x), A @ x, rtol=0, atol=1e-14)
    assert np.allclose(orc.csc_transpose_multiply(P, y), A.T @ y, rtol=0, atol=1e-14)
    assert abs(x @ orc.csc_transpose_multiply(P, y) - orc.csc_multiply(P, x) @ y) < 1e-13


def test_spectral_norm_closed_forms():
    d = cm.dense_lp(np.diag([3.0, -1.0, 0.5]), [0] * 3, [0] * 3, np.zeros(3))
    P = orc.Problem(d)
    assert abs(orc.spectral_norm(P, P.val, 30, 1) - 3.0 * 1.01) < 1e-12
    rng = np.random.default_rng(9)
    A = rng.standard_normal((12, 20))
    A[0] *= 6.0  # a clear gap between the top two singular values
    P = orc.Problem(cm.dense_lp(A, [0] * 12, [0] * 12, np.zeros(20)))
    top = np.linalg.svd(A, compute_uv=False)[0]
    assert abs(orc.spectral_norm(P, P.val, 30, 1) / 1.01 - top) / top < 1e-9


def test_scaling_one_by_one_closed_form():
    P = orc.Problem(cm.dense_lp([[4.0]], [2.0], [2.0], [1.0]))
    s = orc.build_scaling(P)
    assert s["row"][0] == 0.5 and s["column"][0] == 0.5 and s["val"][0] == 1.0
    assert s["row_lower"][0] == 1.0 and s["cost"][0] == 0.5


def test_evaluate_zero_at_a_hand_optimum():
    # min x + 2y s.t. x + y = 4, x, y >= 0: optimum (4, 0); the PDHG multiplier is -1 (the
    # simplex dual negated). Every residual is exactly zero there.
    P = orc.Problem(cm.dense_lp([[1, 1]], [4], [4], [1, 2]))
    act, red, res = orc.evaluate(P, np.array([4.0, 0.0]), np.array([-1.0]))
    assert act.tolist() == [4.0] and red.tolist() == [0.0, 1.0]
    assert res[0] == 0 and res[1] == 0 and res[2] == 0 and res[3] == 4.0 and res[4] == 4.0


@pytest.mark.parametrize("name,make,expected,tol", cm.KNOWN_ANSWERS, ids=[k[0] for k in cm.KNOWN_ANSWERS])
def test_oracle_known_answers(name, make, expected, tol):
    r = orc.solve(orc.Problem(make()), tolerance=1e-8, iteration_limit=200000)
    assert r["status"] == "optimal"
    assert abs(r["objective"] - expected) <= tol


def test_kkt_fixture_plants_the_optimum():
    """The 120 generated instances (tests/golden/kkt_20260904.json): each planted point is
    feasible, and HiGHS's dual simplex (scipy) reaches the planted objective exactly, so the
    construction, not only the draw order, is right."""
    from scipy.optimize import linprog
    inst = cm.kkt_instances()
    assert len(inst) == 120
    dims = set()
    for d, opt in inst:
        P = orc.Problem(d)
        A = np.zeros((P.m, P.n))
        for j in range(P.n):
            A[P.rowidx[P.colptr[j]:P.colptr[j + 1]], j] = P.val[P.colptr[j]:P.colptr[j + 1]]
        dims.add((P.m, P.n))
        r = linprog(P.cost, A_ub=-A, b_ub=-P.row_lower,
                    bounds=[(0, None if math.isinf(u) else u) for u in P.col_upper], method="highs-ds")
        assert r.status == 0 and abs(r.fun - opt) <= 1e-9 * max(1.0, abs(opt)), (r.fun, opt)
    assert min(dims) >= (2, 2) and max(m for m, _ in dims) <= 6 and max(n for _, n in dims) <= 6


def test_oracle_agrees_with_the_planted_optima():
    """tests/unit/test_pdhg.cpp:178-222 on the oracle: at 1e-9, every instance that ends
    optimal is within 1e-6 relative of the optimum, and more than 60 do."""
    compared = 0
    for d, opt in cm.kkt_instances():
        r = orc.solve(orc.Problem(d), tolerance=1e-9, iteration_limit=200000)
        if r["status"] != "optimal":
            continue
        compared += 1
        assert abs(r["objective"] - opt) <= 1e-6 * max(1.0, abs(opt)), (r["objective"], opt)
    assert compared > 60


@pytest.mark.parametrize("name,d,status,objective", cm.DEGENERATE, ids=[c[0] for c in cm.DEGENERATE])
def test_oracle_degenerate_shapes(name, d, status, objective):
    r = orc.solve(orc.Problem(d), tolerance=1e-8, iteration_limit=100000)
    assert r["status"] == status
    if objective is not None:
        assert abs(r["objective"] - objective) <= 1e-6 * max(1.0, abs(objective))


def test_oracle_iteration_limit_and_first_step_guard():
    P = orc.Problem(cm.dense_lp([[1, 1]], [2], [np.inf], [1, 1]))
    r = orc.solve(P, tolerance=1e-12, iteration_limit=20)
    assert r["status"] == "iteration_limit" and r["iterations"] == 20 and math.isinf(r["dual_bound"])
    r = orc.solve(P, tolerance=1e-8, iteration_limit=200000)
    assert r["status"] == "optimal" and r["iterations"] < 500


def test_oracle_blend_guard():
    r = orc.solve(orc.Problem(cm.blend_guard()), tolerance=1e-12, iteration_limit=1000000)
    assert r["status"] == "optimal"
    assert abs(r["objective"] - 214.14594594594595) < 1e-4


@pytest.mark.parametrize("name", ["afiro", "sc50a", "sc50b", "blend", "sc105"])
def test_oracle_netlib_published_optima(name):
    pub = cm.netlib_optima()[name]["published_optimal"]
    r = orc.solve(orc.Problem(cm.netlib(name)), tolerance=1e-8, iteration_limit=1000000)
    assert r["status"] == "optimal"
    assert abs(r["objective"] - pub) / max(1.0, abs(pub)) < 1e-6


def test_netlib_reader_matches_published_dimensions():
    for name, info in cm.netlib_optima().items():
        d = cm.netlib(name)
        assert d["m"] + 1 == info["published_rows"], name
        assert d["n"] == info["published_cols"], name
        assert int(d["colptr"][-1]) + d["cost_nonzeros"] == info["published_nonzeros"], name


# ---- Solution::recompute_quality and Model::validate (src/core/model.cpp) ------------------

def _quality_cases():
    """(name, lp, point) with points that exercise every term: primal violations of both signs
    on bounded and unbounded sides, duals of both signs on equality, one-sided and range rows,
    reduced costs that disagree with c - A^T y, fixed columns, maximise sense and an offset."""
    base = [("C0-s1", cm.c0(1)[1]), ("zipf-2000x5000", cm.generated("zipf", 2000, 5000, 10, 3)[1])]
    for nm in ("afiro", "sc50a", "blend", "adlittle", "share2b", "israel", "stocfor1"):
        base.append((nm, cm.netlib(nm)))
    for nm, lp, _, _ in cm.DEGENERATE:
        base.append((nm, lp))
    out = []
    for k, (nm, d) in enumerate(base):
        rng = np.random.default_rng(100 + k)
        for variant in ("min", "max+offset"):
            lp = dict(d)
            if variant != "min":
                lp = dict(d, maximize=True, objective_offset=3.25)
            P = orc.Problem(lp)
            lo = np.where(np.abs(P.col_lower) < 1e30, P.col_lower, -5.0)
            hi = np.where(np.abs(P.col_upper) < 1e30, P.col_upper, 5.0)
            x = rng.uniform(lo - 0.1, hi + 0.1) if P.n else np.zeros(0)
            if P.n:
                x[rng.random(P.n) < 0.3] = 0.0  # zero columns are skipped by the row scale
            yd = rng.standard_normal(P.m)
            yd[rng.random(P.m) < 0.2] = 0.0
            aty = orc.csc_transpose_multiply(P, yd) if P.m and P.n else np.zeros(P.n)
            dj = P.cost - aty
            if P.n:
                dj[rng.random(P.n) < 0.3] += rng.standard_normal() * 1e-3
            out.append((f"{nm}-{variant}", lp, x, yd, dj, float(rng.normal() * 10)))
    return out


QUALITY_CASES = _quality_cases()


@pytest.mark.skipif(not os.path.exists(orc.REF_SO), reason="oracle/_ref/libref.so not built (no reference tree)")
@pytest.mark.parametrize("case", QUALITY_CASES, ids=[c[0] for c in QUALITY_CASES])
def test_recompute_quality_oracle_and_engine_equal_the_reference(case):
    """or_recompute_quality and the engine's Solution::recompute_quality (sankhya_pdhg_quality)
    against the reference's own compiled Solution::recompute_quality, bit for bit."""
    name, lp, x, yd, dj, db = case
    R = orc.ref()
    P = orc.Problem(lp)
    act_ref, q_ref = np.zeros(max(P.m, 1)), np.zeros(9)
    f = lambda v: np.ascontiguousarray(v if len(v) else np.zeros(1), np.float64)
    R.ref_recompute_quality(P.m, P.n, P.colptr, P.rowidx, f(P.val), f(P.cost), f(P.col_lower),
                            f(P.col_upper), f(P.row_lower), f(P.row_upper), P.offset,
                            1 if P.maximize else 0, f(x), f(yd), f(dj), db, act_ref, q_ref)
    act_or, q_or = orc.recompute_quality(P, f(x), f(yd), f(dj), db)
    assert np.array_equal(act_or, act_ref[: P.m]), name
    assert np.array_equal(q_or, q_ref), (name, q_or, q_ref)
    act_en, q_en = sh.quality(cm.model_of(lp), x, yd, dj, db)
    assert np.array_equal(act_en, act_ref[: P.m]), name
    assert np.array_equal(q_en, q_ref), (name, q_en, q_ref)


def test_quality_cases_violate_every_kind_of_term():
    """The points above are not near-optimal: across the set every field the comparison covers
    is nonzero somewhere, so agreement is not agreement on zeros."""
    seen = np.zeros(9, dtype=bool)
    for name, lp, x, yd, dj, db in QUALITY_CASES:
        _, q = orc.recompute_quality(orc.Problem(lp), x if len(x) else np.zeros(1),
                                     yd if len(yd) else np.zeros(1), dj if len(dj) else np.zeros(1), db)
        seen |= q != 0.0
    assert seen[[0, 1, 2, 3, 4, 5, 7, 8]].all(), seen  # integrality (6) is 0 for an LP


def _validate_cases():
    ok = cm.dense_lp([[1.0, 2.0], [0.0, 3.0]], [1.0, -np.inf], [4.0, 2.0], [1.0, 1.0])
    cases = [("well-formed", ok)]
    def edit(name, **kv):
        lp = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in ok.items()}
        for k, (idx, val) in kv.items():
            lp[k][idx] = val
        cases.append((name, lp))
    edit("nan-entry", val=(1, np.nan))
    edit("infinite-entry", val=(0, np.inf))
    edit("nan-column-bound", col_lower=(1, np.nan))
    edit("column-lower-above-upper", col_lower=(0, 2.5), col_upper=(0, 1.25))
    edit("nan-cost", cost=(1, np.nan))
    edit("nan-row-bound", row_upper=(0, np.nan))
    edit("row-lower-above-upper", row_lower=(1, 7.5), row_upper=(1, 0.5))
    return cases


VALIDATE_CASES = _validate_cases()


@pytest.mark.skipif(not os.path.exists(orc.REF_SO), reason="oracle/_ref/libref.so not built (no reference tree)")
@pytest.mark.parametrize("case", VALIDATE_CASES, ids=[c[0] for c in VALIDATE_CASES])
def test_validate_messages_equal_the_reference(case):
    """The engine's Model::validate (through SANKHYA's C API, sankhya_model_validate) gives the
    reference's message text for each malformed model, and accepts the well-formed one."""
    import ctypes as C
    name, lp = case
    R = orc.ref()
    P = orc.Problem(lp)
    buf = C.create_string_buffer(512)
    R.ref_validate(P.m, P.n, P.colptr, P.rowidx, P.val, P.cost, P.col_lower, P.col_upper,
                   P.row_lower, P.row_upper, buf, 512)
    want = buf.value.decode()
    lib = sh.load_library()
    lib.sankhya_model_validate.argtypes = [C.c_void_p]
    lib.sankhya_last_error.restype = C.c_char_p
    mdl = cm.model_of(lp)
    rc = lib.sankhya_model_validate(mdl._h)
    got = "" if rc == 0 else lib.sankhya_last_error().decode()
    assert got == want, (name, got, want)
    assert (name == "well-formed") == (want == "")


def test_threaded_transpose_gather_repeats_the_single_thread_run():
    """The reference's only parallel loop is the A^T y column gather (src/la/sparse.cpp:244-259),
    one writer per column, so a run is the same at any thread count; the oracle's all-cores
    leg of the CPU baseline relies on that."""
    d = cm.generated("zipf", 2000, 5000, 10, 3)[1]
    P = orc.Problem(d)
    runs = []
    for t in (1, 3, 8):
        orc.set_threads(t)
        runs.append(orc.solve(P, tolerance=1e-6, iteration_limit=400))
    orc.set_threads(1)
    for r in runs[1:]:
        assert r["iterations"] == runs[0]["iterations"]
        assert np.array_equal(r["col_value"], runs[0]["col_value"])
        assert np.array_equal(r["row_dual"], runs[0]["row_dual"])
