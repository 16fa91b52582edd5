"""A differentiable batched LP layer for PyTorch, over the MI355X batched PDHG engine.

Forward: for every LP of a fixed family (matrix, column bounds and row structure fixed),
cost c[k] and row bounds b[k] -> the optimal point x*[k] and row duals y*[k] (the engine's
Solution.col_value and Solution.row_dual, reference sign conventions). b[k][i] sets the finite
side(s) of row i: an equality row becomes rl = ru = b; a one-sided row moves its finite bound
to b; a range row keeps its width and moves its lower bound to b.

Backward: the hand-written adjoint of the engine (implicit differentiation of the PDHG fixed
point on the active set, CGLS on the device; include/pdhg_batch.h, pdhg_batch_adjoint) maps
dL/dx and dL/dy to dL/dc and dL/db. The layer has no CPU path: without the built library and a
gfx950 device it raises.

The layer only moves arrays between torch (float64, CPU) and the C ABI; the solves and the
adjoint run on the GPU.
"""
from __future__ import annotations

import numpy as np
import torch

from . import BatchSession, Model


class LPFamily:
    """The fixed part of a batch of LPs: one Model per LP (its matrix, bounds and sense)."""

    def __init__(self, models, **options):
        self.templates = [m.arrays() for m in models]
        self.options = dict(options)

    @property
    def sizes(self):
        return [(d["m"], d["n"]) for d in self.templates]

    def _model(self, k, c, b):
        d = self.templates[k]
        rl, ru = d["row_lower"].copy(), d["row_upper"].copy()
        fin_l, fin_u = np.abs(rl) < 1e30, np.abs(ru) < 1e30
        width = np.where(fin_l & fin_u, ru - rl, 0.0)
        rl = np.where(fin_l, b, rl)
        ru = np.where(fin_u & fin_l, b + width, np.where(fin_u, b, ru))
        return Model.from_csc(d["m"], d["n"], d["colptr"], d["rowidx"], d["val"], c,
                              d["col_lower"], d["col_upper"], rl, ru, d["objective_offset"],
                              d["maximize"])

    def row_bound(self, k):
        """The b that reproduces LP k's own row bounds."""
        d = self.templates[k]
        rl, ru = d["row_lower"], d["row_upper"]
        return np.where(np.abs(rl) < 1e30, rl, np.where(np.abs(ru) < 1e30, ru, 0.0))


class _LPSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, b, family, tol, max_iterations):
        cs = c.detach().cpu().numpy().astype(np.float64)
        bs = b.detach().cpu().numpy().astype(np.float64)
        models = [family._model(k, cs[k], bs[k]) for k in range(cs.shape[0])]
        session = BatchSession(models, **family.options)
        tolerance = float(family.options.get("pdhg_tolerance", 1e-8))
        bad = [s.status for s in session.solutions if not s.reached(tolerance)]
        if bad:
            session.close()
            raise RuntimeError(f"LP layer: {len(bad)} LPs not solved ({bad[0]})")
        x = torch.tensor(np.stack([s.col_value for s in session.solutions]), dtype=torch.float64)
        y = torch.tensor(np.stack([s.row_dual for s in session.solutions]), dtype=torch.float64)
        ctx.session = session
        ctx.tol, ctx.max_iterations = tol, max_iterations
        ctx.models = models  # the session refers to them
        return x, y

    @staticmethod
    def backward(ctx, gx, gy):
        session = ctx.session
        gxs = None if gx is None else [v.numpy() for v in gx.detach().cpu().to(torch.float64)]
        gys = None if gy is None else [v.numpy() for v in gy.detach().cpu().to(torch.float64)]
        gc, gb, its, res = session.adjoint(gxs, gys, ctx.tol, ctx.max_iterations)
        ctx.adjoint_iterations, ctx.adjoint_residuals = its, res
        return (torch.tensor(np.stack(gc), dtype=torch.float64),
                torch.tensor(np.stack(gb), dtype=torch.float64), None, None, None)


def lp_layer(c: torch.Tensor, b: torch.Tensor, family: LPFamily, tol: float = 1e-14,
             max_iterations: int = 50000):
    """(x*, y*) = argmin over the family's LPs with costs c (batch x n) and row bounds b
    (batch x m); differentiable in c and b."""
    return _LPSolve.apply(c, b, family, tol, max_iterations)
