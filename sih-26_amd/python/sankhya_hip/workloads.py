"""Synthetic workloads of the benchmark and the tests, built by the engine's own deterministic
generator (sankhya_model_generate).

  batched: LP i has m_i rows, m_i = round(exp(U[ln 200, ln 20000])) from numpy's default_rng(seed),
           n_i = round(2.5 m_i), 6 nonzeros per column at uniform rows, equality rows b = A x0,
           c = A^T y0 + z (the 1,000 x 3,000 instance's construction), generator seed seed + i;
  blend:   m equality rows (one dense, sum x = sum x0), n columns, 5 nonzeros per column,
           40% of the columns upper-bounded by U[1,50], c ~ U[-5,5], seed seed0 + i.
"""
from __future__ import annotations

import numpy as np

from . import Model


def batched_sizes(count: int, seed: int = 1000):
    rng = np.random.default_rng(seed)
    m = np.rint(np.exp(rng.uniform(np.log(200.0), np.log(20000.0), count))).astype(int)
    return [(int(mi), int(round(2.5 * mi))) for mi in m]


def batched(count: int, seed: int = 1000):
    return [Model.generate("uniform", m, n, 6, seed + i)
            for i, (m, n) in enumerate(batched_sizes(count, seed))]


def blend(count: int, m: int = 5000, n: int = 12000, seed0: int = 5000):
    return [Model.generate("blend", m, n, 5, seed0 + i) for i in range(count)]
