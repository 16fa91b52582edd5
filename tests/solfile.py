"""Reader of the solution file format (header `key value` lines; `begin columns n` ...
`end columns` with name, value, reduced cost, basis status; `begin rows m` ... `end rows` with
name, activity, dual, basis status; names with whitespace in double quotes with `"` and `\\`
escaped), written from the format as the reference's own checker reads it
(tools/verify_solution_sol.py:57-137). It shares no code with the engine's writer, so a
disagreement between the two surfaces in the tests. Test infrastructure.
"""
from __future__ import annotations


def split_record(line: str) -> list:
    if not line.startswith('"'):
        return line.split()
    name, i = [], 1
    while i < len(line):
        ch = line[i]
        if ch == "\\" and i + 1 < len(line):
            name.append(line[i + 1])
            i += 2
            continue
        if ch == '"':
            return ["".join(name)] + line[i + 1:].split()
        name.append(ch)
        i += 1
    return line.split()  # no closing quote: keep every field rather than truncate


def read_sol(path) -> dict:
    """{'header': {key: text}, 'columns': {name: (value, reduced, status)},
    'rows': {name: (activity, dual, status)}, 'column_order': [...], 'row_order': [...],
    'declared': {block: count}}."""
    out = {"header": {}, "columns": {}, "rows": {}, "column_order": [], "row_order": [],
           "declared": {}}
    block = None
    with open(path, encoding="utf-8") as fh:
        for raw in fh:
            line = raw.strip()
            if not line or line.startswith("#"):
                continue
            f = split_record(line)
            if f[0] == "begin":
                block = f[1]
                out["declared"][block] = int(f[2]) if len(f) > 2 else None
                continue
            if f[0] == "end":
                block = None
                continue
            if block in ("columns", "rows"):
                rec = (float(f[1]), float(f[2]), f[3] if len(f) > 3 else "unknown")
                out[block][f[0]] = rec
                out["column_order" if block == "columns" else "row_order"].append(f[0])
            elif block is None and len(f) >= 2:
                out["header"][f[0]] = " ".join(f[1:])
    return out
