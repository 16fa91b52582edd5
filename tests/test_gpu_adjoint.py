"""The adjoint of the batched engine (pdhg_batch_adjoint) and the torch layer over it, on a
gfx950 device.

Workload: blend-style LPs (m = 5,000 equality rows, one of them dense, n = 12,000 columns, 5
nonzeros per column plus the dense row, 40% of the columns with upper bounds U[1,50],
c ~ U[-5,5]; seeds 5000 + i). Checks:
  - the envelope identities through torch autograd: for L = c.x*, dL/dc = x* and
    dL/db = the row duals, within 1e-5 relative of their largest entry, on every LP whose
    active system is square; on a degenerate vertex (a basic column at its bound, fewer free
    columns than rows) the duals are not unique, and dL/db must instead be an exact dual of
    the active set, A_F^T (dL/db) = c_F;
  - the device CGLS answers against direct sparse-LU solves on the solution's active set (scipy,
    CPU): dL/db for L = w.x* and dL/dc for L = v.row_dual, every entry within 1e-8 of the
    gradient's largest entry; and central finite differences (h = 1e-6) of those exact solution
    maps on random coordinates within 1e-5 of it (at h = 1e-6 the rounding of each solve,
    amplified by 1/h, sets that floor); dL/dc = 0 for L = w.x*;
  - dL/db for L = c.x* against the exact row duals of a simplex vertex (HiGHS, as shipped in
    scipy) on small blend LPs (500 x 1,200, solved to 1e-8) with square active systems, within
    1e-8, on at least three of four;
  - central finite differences (h = 1e-6) of re-solved LPs for three coordinates of b: each
    perturbed LP is solved again by PDHG, its active set read from its own solution, and its
    vertex taken exactly on that active set; where the active set is the unperturbed one (at
    least one coordinate must be), the difference matches the adjoint within 1e-5 of the
    gradient's largest entry.
    This checks the active set the adjoint uses against re-solves.
"""
import time

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import common as cm
import sankhya_hip as sh
import sankhya_hip.workloads as wl

torch = pytest.importorskip("torch")
from sankhya_hip.torch_layer import LPFamily, lp_layer  # noqa: E402

pytestmark = pytest.mark.gpu

OPTS = dict(pdhg_tolerance=1e-9, pdhg_stop_rule="relative_kkt", iteration_limit=3_000_000,
            pdhg_polish=False)


@pytest.fixture(scope="module", autouse=True)
def device():
    desc = cm.gpu_or_skip()
    print(f"[gpu] {desc}", flush=True)


def blend(count, m=5000, n=12000, seed0=5000):
    return wl.blend(count, m, n, seed0)


def _reduced_cost(d, row_dual):
    """c - A^T row_dual: the engine's col_dual for a minimise model."""
    A = sp.csc_matrix((d["val"], d["rowidx"], d["colptr"]), shape=(d["m"], d["n"]))
    return d["cost"] - A.T @ row_dual


def free_columns(d, x, reduced):
    """The device's complementarity rule: free when the distance to the nearest bound exceeds
    the reduced cost."""
    lo, hi = d["col_lower"], d["col_upper"]
    with np.errstate(invalid="ignore"):
        slack = np.minimum(np.where(np.abs(lo) < 1e30, x - lo, np.inf),
                           np.where(np.abs(hi) < 1e30, hi - x, np.inf))
    return np.where(slack > np.abs(reduced))[0]


def active_maps(d, sol):
    """Exact solution maps of one LP on its active set: x(b) and row_dual(c) (min sense, all
    rows active as in the blend family)."""
    A = sp.csc_matrix((d["val"], d["rowidx"], d["colptr"]), shape=(d["m"], d["n"]))
    x = sol.col_value
    lo, hi = d["col_lower"], d["col_upper"]
    with np.errstate(invalid="ignore"):
        slack = np.minimum(np.where(np.abs(lo) < 1e30, x - lo, np.inf),
                           np.where(np.abs(hi) < 1e30, hi - x, np.inf))
    # the same complementarity rule as the device: free when the slack exceeds |c + A^T y|
    at = ~(slack > np.abs(sol.col_dual))
    F = np.where(~at)[0]
    N = np.where(at)[0]
    assert len(F) == d["m"], f"degenerate vertex: {len(F)} free columns for {d['m']} rows"
    M = A[:, F].tocsc()
    lu = spla.splu(M)
    AN_xN = A[:, N] @ x[N]

    def x_of(b):
        out = x.copy()
        out[F] = lu.solve(b - AN_xN)
        return out

    def rowdual_of(c):
        return lu.solve(c[F], trans="T")

    def vjp_b(w):  # d(w.x*)/db on the active set, directly: M^-T w_F
        return lu.solve(w[F], trans="T")

    def vjp_c(v):  # d(v.row_dual)/dc on the active set, directly: M^-1 v, on F
        out = np.zeros(d["n"])
        out[F] = lu.solve(v)
        return out

    x_of.vjp_b, x_of.vjp_c = vjp_b, vjp_c
    return x_of, rowdual_of, F


def test_envelope_gradients_through_the_torch_layer():
    models = blend(4)
    fam = LPFamily(models, **OPTS)
    c = torch.tensor(np.stack([t["cost"] for t in fam.templates]), requires_grad=True)
    b = torch.tensor(np.stack([fam.row_bound(k) for k in range(4)]), requires_grad=True)
    t = time.time()
    x, y = lp_layer(c, b, fam)
    tf = time.time() - t
    obj = (c * x).sum()
    t = time.time()
    obj.backward()
    tb = time.time() - t
    print(f"forward (4 LPs) {tf:.2f}s, backward {tb:.3f}s", flush=True)
    square = 0
    for k in range(4):
        xk, yk = x[k].detach().numpy(), y[k].detach().numpy()
        ex = np.abs(c.grad[k].numpy() - xk).max() / max(1.0, np.abs(xk).max())
        ey = np.abs(b.grad[k].numpy() - yk).max() / max(1.0, np.abs(yk).max())
        d = fam.templates[k]
        F = free_columns(d, xk, _reduced_cost(d, yk))
        print(f"  LP {k}: |F| {len(F)} for {d['m']} rows; |dL/dc - x*| {ex:.1e}, |dL/db - y*| {ey:.1e}",
              flush=True)
        assert ex <= 1e-5
        # an empty row (possible in the generator's random pattern) is 0 = 0 and leaves A one
        # rank short; the active system is square when F matches the rows that have entries
        if len(F) == int(np.count_nonzero(np.bincount(d["rowidx"], minlength=d["m"]))):
            square += 1
            assert ey <= 1e-5
        else:
            # a degenerate vertex (a basic column at its bound): the row duals are not unique
            # and dL/db is a subgradient; the adjoint's is still an exact dual of the active
            # set, A_F^T (dL/db) = c_F
            # (checked on the columns that are free beyond doubt, slack > 10 |reduced cost|: at
            # a degenerate vertex the tie column can read either way in the last digits)
            A = sp.csc_matrix((d["val"], d["rowidx"], d["colptr"]), shape=(d["m"], d["n"]))
            Fs = free_columns(d, xk, 10.0 * _reduced_cost(d, yk))
            c_f = d["cost"][Fs]
            resid = np.abs(A[:, Fs].T @ b.grad[k].numpy() - c_f).max() / np.abs(c_f).max()
            print(f"    degenerate: active-system residual {resid:.1e} on {len(Fs)} columns", flush=True)
            assert resid <= 1e-8
    assert square >= 3


def test_envelope_gradient_against_simplex_duals():
    """dL/db for L = c.x* against the exact row duals of a simplex vertex (HiGHS through scipy)
    on small blend LPs (m = 500, n = 1,200): the adjoint solves the active system exactly, so
    it matches the vertex duals far more closely than PDHG's own y* does."""
    models = blend(4, m=500, n=1200)
    sess = sh.BatchSession(models, **dict(OPTS, pdhg_tolerance=1e-8))
    datas = [md.arrays() for md in models]
    gc, gb, its, res = sess.adjoint([d["cost"] for d in datas], None)
    square = 0
    for k, (d, s) in enumerate(zip(datas, sess.solutions)):
        if not s.reached(1e-8):
            print(f"  LP {k}: {s.status}, not checked", flush=True)
            continue
        r, eq = cm.simplex_result(d)
        exact = np.zeros(d["m"])
        exact[eq] = r.eqlin.marginals
        F = free_columns(d, s.col_value, s.col_dual)
        rows = int(np.count_nonzero(np.bincount(d["rowidx"], minlength=d["m"])))
        scale = max(1.0, np.abs(exact).max())
        err = np.abs(gb[k] - exact).max() / scale
        err_pdhg = np.abs(s.row_dual - exact).max() / scale
        print(f"  LP {k}: |F| {len(F)} for {rows} rows with entries; adjoint vs simplex duals {err:.1e}, "
              f"PDHG y* vs simplex duals {err_pdhg:.1e}", flush=True)
        if len(F) == rows:
            square += 1
            assert err <= 1e-8
    assert square >= 3


def test_adjoint_against_finite_differences_of_the_exact_maps():
    models = blend(2, seed0=5100)
    sess = sh.BatchSession(models, **OPTS)
    rng = np.random.default_rng(0)
    datas = [md.arrays() for md in models]
    w = [rng.standard_normal(d["n"]) for d in datas]
    v = [rng.standard_normal(d["m"]) for d in datas]
    gc_x, gb_x, its_x, res_x = sess.adjoint(w, None)
    gc_y, gb_y, its_y, res_y = sess.adjoint(None, v)
    print(f"CGLS iterations x {its_x[:, 0]}, y {its_y[:, 1]}; residuals {res_x[:, 0]}, {res_y[:, 1]}", flush=True)
    h = 1e-6
    for k, (d, s) in enumerate(zip(datas, sess.solutions)):
        assert s.reached(OPTS["pdhg_tolerance"])
        x_of, rowdual_of, F = active_maps(d, s)
        b0 = d["row_upper"].copy()
        c0 = d["cost"].copy()
        scale_b = np.abs(gb_x[k]).max()
        direct_b = x_of.vjp_b(w[k])
        err_b = np.abs(gb_x[k] - direct_b).max() / scale_b
        direct_c = x_of.vjp_c(v[k])
        err_c = np.abs(gc_y[k] - direct_c).max() / np.abs(direct_c).max()
        print(f"  LP {k}: adjoint vs direct sparse LU: dL/db {err_b:.1e}, dL/dc {err_c:.1e}", flush=True)
        assert err_b <= 1e-8 and err_c <= 1e-8
        for i in rng.choice(d["m"], 6, replace=False):
            e = np.zeros(d["m"])
            e[i] = h
            fd = (w[k] @ x_of(b0 + e) - w[k] @ x_of(b0 - e)) / (2 * h)
            assert abs(fd - gb_x[k][i]) <= 1e-5 * scale_b, (k, i, fd, gb_x[k][i])
        assert np.abs(gc_x[k]).max() <= 1e-9 * max(1.0, scale_b)  # x* does not move with c
        scale_c = np.abs(gc_y[k]).max()
        for j in rng.choice(F, 6, replace=False):
            e = np.zeros(d["n"])
            e[j] = h
            fd = (v[k] @ rowdual_of(c0 + e) - v[k] @ rowdual_of(c0 - e)) / (2 * h)
            assert abs(fd - gc_y[k][j]) <= 1e-5 * scale_c, (k, j, fd, gc_y[k][j])
        assert np.abs(gb_y[k]).max() == 0.0  # row duals do not move with b on the active set
    sess.close()


def test_adjoint_against_finite_differences_of_resolves():
    md = blend(1, seed0=5200)[0]
    d = md.arrays()
    fam = LPFamily([md], **OPTS)
    sess = sh.BatchSession([md], **OPTS)
    rng = np.random.default_rng(1)
    w = rng.standard_normal(d["n"])
    _, gb, _, _ = sess.adjoint([w], None)
    base = sess.solutions[0]
    sess.close()
    _, _, F0 = active_maps(d, base)
    b0 = fam.row_bound(0)
    idx = rng.choice(d["m"] - 1, 3, replace=False)
    h = 1e-6
    models = []
    for i in idx:
        for sgn in (1.0, -1.0):
            bb = b0.copy()
            bb[i] += sgn * h
            models.append((bb, fam._model(0, d["cost"], bb)))
    sols = sh.solve_batch([m for _, m in models], **OPTS)
    stable = 0
    for q, i in enumerate(idx):
        vals = []
        for side in (0, 1):
            bb, md_s = models[2 * q + side]
            s = sols[2 * q + side]
            x_of, _, F = active_maps(md_s.arrays(), s)
            if not np.array_equal(F, F0):
                break
            vals.append(w @ x_of(bb))  # the re-solved LP's own vertex, exact on its active set
        if len(vals) < 2:
            print(f"  b[{i}]: active set changes within {h:g}; skipped", flush=True)
            continue
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"  b[{i}]: re-solved FD {fd:.10g}, adjoint {gb[0][i]:.10g}", flush=True)
        assert abs(fd - gb[0][i]) <= 1e-5 * np.abs(gb[0]).max()
        stable += 1
    assert stable >= 1
