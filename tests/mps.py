"""A small MPS reader for the committed Netlib fixtures (test infrastructure).

Covers fixed and free MPS: NAME, ROWS (N/E/L/G; the first N row is the objective, others are
dropped), COLUMNS (integer MARKER lines ignored), RHS (an RHS on the objective row is the
negated objective constant), RANGES (IBM semantics: E rows widen by |R| up when R > 0, down
when R < 0; L rows get [rhs-|R|, rhs]; G rows [rhs, rhs+|R|]) and BOUNDS (UP, LO, FX, FR, MI,
PL, BV; an UP bound below zero on a column whose lower bound is still 0 makes the lower bound
-inf). Values of magnitude >= 1e30 are infinities. The result is CSC with rows ascending per
column, as the engine takes it.
"""
from __future__ import annotations

import math
from typing import Dict

import numpy as np

INF = math.inf


def _num(tok: str) -> float:
    v = float(tok)
    if abs(v) >= 1e30:
        return INF if v > 0 else -INF
    return v


def read_mps(path: str) -> Dict:
    section = None
    row_type: Dict[str, str] = {}
    row_index: Dict[str, int] = {}
    obj_name = None
    cols: Dict[str, Dict[int, float]] = {}
    col_order = []
    cost: Dict[str, float] = {}
    rhs: Dict[int, float] = {}
    ranges: Dict[int, float] = {}
    offset = 0.0
    lower: Dict[str, float] = {}
    upper: Dict[str, float] = {}
    name = ""
    with open(path) as fh:
        for raw in fh:
            line = raw.rstrip("\n")
            if not line.strip() or line.startswith("*"):
                continue
            if not line[0].isspace():
                parts = line.split()
                section = parts[0].upper()
                if section == "NAME" and len(parts) > 1:
                    name = parts[1]
                if section == "ENDATA":
                    break
                continue
            tok = line.split()
            if section == "ROWS":
                kind, rname = tok[0].upper(), tok[1]
                if kind == "N":
                    if obj_name is None:
                        obj_name = rname
                    continue
                row_type[rname] = kind
                row_index[rname] = len(row_index)
            elif section == "COLUMNS":
                if len(tok) >= 3 and tok[1].strip("'").upper() == "MARKER":
                    continue
                cname = tok[0]
                if cname not in cols:
                    cols[cname] = {}
                    col_order.append(cname)
                for k in range(1, len(tok) - 1, 2):
                    rname, v = tok[k], float(tok[k + 1])
                    if rname == obj_name:
                        cost[cname] = cost.get(cname, 0.0) + v
                    elif rname in row_index:
                        i = row_index[rname]
                        cols[cname][i] = cols[cname].get(i, 0.0) + v
            elif section == "RHS":
                start = 1 if len(tok) % 2 == 1 else 0
                for k in range(start, len(tok) - 1, 2):
                    rname, v = tok[k], float(tok[k + 1])
                    if rname == obj_name:
                        offset = -v
                    elif rname in row_index:
                        rhs[row_index[rname]] = v
            elif section == "RANGES":
                start = 1 if len(tok) % 2 == 1 else 0
                for k in range(start, len(tok) - 1, 2):
                    rname, v = tok[k], float(tok[k + 1])
                    if rname in row_index:
                        ranges[row_index[rname]] = v
            elif section == "BOUNDS":
                btype = tok[0].upper()
                if btype in ("FR", "MI", "PL", "BV"):
                    cname = tok[2] if len(tok) >= 3 else tok[1]
                    v = None
                else:
                    cname, v = tok[2], _num(tok[3])
                if btype == "UP":
                    upper[cname] = v
                    if v < 0 and lower.get(cname, 0.0) == 0.0:
                        lower[cname] = -INF
                elif btype == "LO":
                    lower[cname] = v
                elif btype == "FX":
                    lower[cname] = v
                    upper[cname] = v
                elif btype == "FR":
                    lower[cname] = -INF
                    upper[cname] = INF
                elif btype == "MI":
                    lower[cname] = -INF
                elif btype == "PL":
                    upper[cname] = INF
                elif btype == "BV":
                    lower[cname] = 0.0
                    upper[cname] = 1.0
    m, n = len(row_index), len(col_order)
    rl, ru = np.zeros(m), np.zeros(m)
    for rname, i in row_index.items():
        b = rhs.get(i, 0.0)
        t = row_type[rname]
        if t == "E":
            rl[i] = ru[i] = b
        elif t == "L":
            rl[i], ru[i] = -INF, b
        elif t == "G":
            rl[i], ru[i] = b, INF
        if i in ranges:
            r = ranges[i]
            if t == "E":
                if r > 0:
                    ru[i] = b + abs(r)
                else:
                    rl[i] = b - abs(r)
            elif t == "L":
                rl[i] = b - abs(r)
            elif t == "G":
                ru[i] = b + abs(r)
    colptr, rowidx, val = [0], [], []
    c = np.zeros(n)
    cl, cu = np.zeros(n), np.full(n, INF)
    for j, cname in enumerate(col_order):
        for i in sorted(cols[cname]):
            v = cols[cname][i]
            if v != 0.0:
                rowidx.append(i)
                val.append(v)
        colptr.append(len(rowidx))
        c[j] = cost.get(cname, 0.0)
        cl[j] = lower.get(cname, 0.0)
        cu[j] = upper.get(cname, INF)
    return dict(name=name, m=m, n=n, colptr=np.array(colptr, np.int32),
                rowidx=np.array(rowidx, np.int32), val=np.array(val, np.float64), cost=c,
                col_lower=cl, col_upper=cu, row_lower=rl, row_upper=ru, objective_offset=offset,
                maximize=False, cost_nonzeros=int(np.count_nonzero(c)),
                row_names=sorted(row_index, key=row_index.get), col_names=list(col_order))
