"""Writes tests/golden/ref_vectors.npz: outputs of the REFERENCE's own build_scaling,
estimate_spectral_norm and pdhg::evaluate (src/la/scaling.cpp, src/pdhg/pdhg_evaluate.cpp),
compiled in place from the reference sources by `make -C oracle ref`, on small inputs that the
tests rebuild (Netlib fixtures and seeded generated instances). The inputs are stored beside the
outputs. Run in the container that has the reference tree:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "sih-26_amd", "python")]

import oracle as orc  # noqa: E402
import common as cm  # noqa: E402


def cases():
    yield "afiro", cm.netlib("afiro")
    yield "blend", cm.netlib("blend")
    yield "uniform_100x300_s5", cm.generated("uniform", 100, 300, 5, 5)[1]
    yield "zipf_300x800_s4", cm.generated("zipf", 300, 800, 10, 4)[1]


def main():
    R = orc.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    out = {}
    for name, d in cases():
        P = orc.Problem(d)
        s = {k: np.zeros(max(n, 1)) for k, n in (("row", P.m), ("column", P.n), ("val", P.nnz),
                                                  ("cost", P.n), ("l", P.n), ("u", P.n),
                                                  ("rl", P.m), ("ru", P.m))}
        R.ref_build_scaling(P.m, P.n, P.colptr, P.rowidx, P.val, P.cost, P.col_lower, P.col_upper,
                            P.row_lower, P.row_upper, P.min_cost, 10, s["row"], s["column"],
                            s["val"], s["cost"], s["l"], s["u"], s["rl"], s["ru"])
        norm = R.ref_spectral_norm(P.m, P.n, P.colptr, P.rowidx, s["val"], 30, 1)
        rng = np.random.default_rng(11)
        x = np.ascontiguousarray(np.clip(rng.uniform(-1, 12, P.n), P.col_lower, np.minimum(P.col_upper, 1e6)))
        y = np.ascontiguousarray(rng.standard_normal(P.m))
        act, red, res = np.zeros(max(P.m, 1)), np.zeros(max(P.n, 1)), np.zeros(9)
        R.ref_evaluate(P.m, P.n, P.colptr, P.rowidx, P.val, P.cost, P.col_lower, P.col_upper,
                       P.row_lower, P.row_upper, P.min_cost, P.bound_norm, P.cost_norm, x, y, act,
                       red, res)
        for k in ("colptr", "rowidx", "val", "cost", "col_lower", "col_upper", "row_lower", "row_upper"):
            out[f"{name}/in_{k}"] = getattr(P, k)
        out[f"{name}/in_dims"] = np.array([P.m, P.n, P.nnz, int(P.maximize)])
        out[f"{name}/in_offset"] = np.array([P.offset])
        for k in ("row", "column", "val", "cost", "l", "u", "rl", "ru"):
            n = {"row": P.m, "rl": P.m, "ru": P.m, "val": P.nnz}.get(k, P.n)
            out[f"{name}/scaling_{k}"] = s[k][:n]
        out[f"{name}/spectral_norm_seed1"] = np.array([norm])
        out[f"{name}/eval_x"], out[f"{name}/eval_y"] = x, y
        out[f"{name}/eval_activity"], out[f"{name}/eval_reduced"] = act[: P.m], red[: P.n]
        out[f"{name}/eval_residuals"] = res
        print(name, P.m, P.n, P.nnz, norm)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)


if __name__ == "__main__":
    main()
