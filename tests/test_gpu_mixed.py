"""Mixed precision on a gfx950 device: FP32 matrix values and iterates in the attempt kernels,
FP64 residuals, restart and primal-weight decisions (the evaluation ticks read the FP64
matrix), and the FP64 continuation ("polish") that finishes the same run.

The reference is FP64 only, so parity here is against the engine's own FP64 path and the
oracle on the same inputs:
  - C0 (1,000 x 3,000) mixed with the polish to 1e-8 (stop at request): objective within 1e-7
    relative of the FP64 solve, of the oracle and of the FP64 simplex optimum (HiGHS, as
    shipped in scipy), and the point accepted by tests/verify.py;
  - the 1M x 2M Zipf LP (nnz 20M) mixed to 1e-6 relative KKT: same status as the FP64 solve,
    objective within 1e-5 relative (both are 1e-6-accurate points; the difference printed is
    the error the mixed path reports against FP64);
  - the handoff itself through the C ABI: the iteration count carries on, nothing restarts.
"""
import time

import numpy as np
import pytest

import common as cm
import oracle as orc
import sankhya_hip as sh
from verify import accepted

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    desc = cm.gpu_or_skip()
    print(f"[gpu] {desc}", flush=True)


def _sol(s):
    return dict(status=s.status, objective=s.objective, col_value=s.col_value, row_dual=s.row_dual,
                col_dual=s.col_dual, row_activity=s.row_activity)


def test_c0_mixed_polished_to_1e8_matches_fp64_and_oracle():
    mdl, d = cm.c0(1)
    opts = dict(pdhg_tolerance=1e-8, pdhg_stop_at_request=True)
    t = time.time()
    a = sh.solve(mdl, **opts)
    b = sh.solve(mdl, pdhg_precision="mixed", **opts)
    print(f"fp64 it={a.iterations} obj={a.objective:.12g}; mixed it={b.iterations} "
          f"(fp32 {b.stats['fp32_iterations']}) obj={b.objective:.12g} ({time.time()-t:.1f}s)", flush=True)
    assert b.algorithm == "pdhg-gpu-mixed" and b.stats["fp32_iterations"] > 0
    assert b.status == a.status and b.status in ("optimal", "feasible")
    assert abs(a.objective - b.objective) <= 1e-7 * max(1.0, abs(a.objective))
    r = orc.solve(orc.Problem(d), tolerance=1e-8, stop_at_request=True)
    assert abs(b.objective - r["objective"]) <= 1e-7 * max(1.0, abs(r["objective"]))
    ok, checks = accepted(d, _sol(b))
    assert ok, checks
    # against the FP64 simplex optimum (HiGHS through scipy: an exact vertex)
    exact = cm.simplex_objective(d)
    err_mixed = abs(b.objective - exact) / max(1.0, abs(exact))
    err_fp64 = abs(a.objective - exact) / max(1.0, abs(exact))
    print(f"simplex optimum {exact:.15g}: mixed error {err_mixed:.1e}, fp64 error {err_fp64:.1e}", flush=True)
    assert err_mixed <= 1e-7 and err_fp64 <= 1e-7


def test_zipf_1m_2m_mixed_against_fp64():
    mdl = sh.Model.generate("zipf", 1_000_000, 2_000_000, 10, 3)
    opts = dict(pdhg_tolerance=1e-6, pdhg_stop_rule="relative_kkt", iteration_limit=300000)
    t = time.time()
    a = sh.solve(mdl, **opts)
    ta = time.time() - t
    t = time.time()
    b = sh.solve(mdl, pdhg_precision="mixed", **opts)
    tb = time.time() - t
    err = abs(a.objective - b.objective) / max(1.0, abs(a.objective))
    print(f"fp64 {a.status} it={a.iterations} {ta:.2f}s obj={a.objective:.12g}; mixed {b.status} "
          f"it={b.iterations} (fp32 {b.stats['fp32_iterations']}) {tb:.2f}s obj={b.objective:.12g}; "
          f"relative objective error vs fp64 {err:.2e}", flush=True)
    assert a.reached(1e-6) and b.reached(1e-6) and a.status == b.status
    assert err <= 1e-5
    for s in (a, b):
        assert max(s.stats["final_primal"], s.stats["final_dual"], s.stats["final_gap"]) <= 1e-6


def test_dense_columns_mixed_against_fp64():
    """Mixed precision on an LP with nearly full columns (FP32 values on the column side's
    long-row segments and tail): polished to 1e-8 under the reference's rule, the same status as
    the FP64 engine and the objective within 1e-7 relative."""
    mdl, d = cm.transposed_zipf(2000, 5000, 3)
    opts = dict(pdhg_tolerance=1e-8, iteration_limit=1000000)
    a = sh.solve(mdl, pdhg_precision="fp64", **opts)
    b = sh.solve(mdl, pdhg_precision="mixed", **opts)
    print(f"dense columns: fp64 {a.status} {a.objective:.12g} it={a.iterations}; mixed {b.status} "
          f"{b.objective:.12g} it={b.iterations} (fp32 {b.stats['fp32_iterations']})", flush=True)
    assert a.status == b.status, (a.message, b.message)
    assert abs(a.objective - b.objective) <= 1e-7 * max(1.0, abs(a.objective))
    assert b.stats["fp32_iterations"] > 0


def test_promote_continues_the_same_run():
    mdl, _ = cm.c0(2)
    norm = sh.solve(mdl, iteration_limit=1).stats["spectral_norm"]
    dev = sh.Device(mdl, precision="mixed")
    dev.start(tolerance=1e-4, spectral_norm=norm, stop_rule=1, iteration_limit=1000000)
    st = dev.run(1)
    while not (st.done or st.limit_hit):
        st = dev.run(8)
    assert st.done
    it0, r0 = st.iteration, st.restarts
    dev.promote(tolerance=1e-7, spectral_norm=norm, stop_rule=1, iteration_limit=1000000)
    st = dev.run(1)
    assert st.iteration >= it0 and st.restarts >= r0
    while not (st.done or st.limit_hit):
        st = dev.run(8)
    assert st.done and st.best_worst <= 1e-7
    with pytest.raises(RuntimeError):
        dev.promote(tolerance=1e-8, spectral_norm=norm, stop_rule=1)  # already FP64
    x, y, red, res = dev.finish()
    assert max(res[0], res[1], res[2]) <= 1e-7
    dev.close()
