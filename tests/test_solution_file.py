"""The solution file (sankhya_write_solution, the reference's .sol format with `certificate
none`), after the reference's own writer tests (tests/unit/test_writer.cpp:107-372): every
number reads back bit for bit, the header carries the fields the checker needs, no negative
zero and no 1e30 stand-in for infinity is printed, an unwritable path is reported, names with
whitespace are quoted and recovered, ordinary names are not quoted, and a status that claims no
point writes no point.

The writer needs no GPU, so these run on the CPU with hand-made solutions; the last test solves
with the engine on a gfx950 device, writes the file, and runs the acceptance check on what the
file alone says.
"""
import math

import numpy as np
import pytest

import common as cm
import sankhya_hip as sh
from solfile import read_sol
from verify import accepted

INF = math.inf


def writertest(col_names=("AL", "BN", "MU"), row_names=("THRUPUT", "DIESEL", "SULPHUR")):
    """The reference test's 3 x 3 maximisation (test_writer.cpp:109-140)."""
    rows = [[1.0, 1.0, 1.0], [0.30, 0.45, 0.38], [0.80, -0.86, -0.22]]
    m = sh.Model.from_dense(rows, [90.0, 40.0, -INF], [120.0, 40.0, 0.0], [2.4, 1.6, 1.64],
                            col_lower=[10.0, 0.0, 0.0], col_upper=[INF, 45.0, 60.0],
                            objective_offset=2.5, maximize=True)
    m.set_names("WRITERTEST", "in-memory", list(col_names), list(row_names))
    return m


def made_solution(status="optimal", **kw):
    vals = dict(status=status, objective=0.1 + 0.2, dual_bound=1.0 / 3.0, iterations=1234,
                col_value=np.array([10.0 / 3.0, math.pi, 1e-300]),
                row_dual=np.array([-0.0, 2.0 / 7.0, -1.5e-17]),
                col_dual=np.array([5e-324, -0.0, 123456789.123456789]),
                row_activity=np.array([97.15999999999998, 40.0, -1e-16]),
                algorithm="pdhg-gpu", message="", solve_seconds=0.125,
                primal_infeasibility=3.0e-9, dual_infeasibility=0.0,
                complementarity_violation=0.0, relative_gap=0.0)
    vals.update(kw)
    return sh.Solution(**vals)


def test_every_number_round_trips_bit_for_bit(tmp_path):
    m, s = writertest(), made_solution()
    path = tmp_path / "out.sol"
    sh.write_solution(m, s, path)
    f = read_sol(path)
    assert f["declared"] == {"columns": 3, "rows": 3}
    assert f["column_order"] == ["AL", "BN", "MU"] and f["row_order"] == ["THRUPUT", "DIESEL", "SULPHUR"]
    for j, name in enumerate(f["column_order"]):
        v, d, st = f["columns"][name]
        assert v == s.col_value[j] and d == s.col_dual[j] and st == "unknown"
    for i, name in enumerate(f["row_order"]):
        a, y, st = f["rows"][name]
        assert a == s.row_activity[i] and y == s.row_dual[i] and st == "unknown"
    h = f["header"]
    assert float(h["objective"]) == s.objective and float(h["dual_bound"]) == s.dual_bound
    assert float(h["solve_seconds"]) == 0.125 and float(h["primal_infeasibility"]) == 3.0e-9


def test_header_carries_the_fields_the_checker_needs(tmp_path):
    m = writertest()
    path = tmp_path / "h.sol"
    sh.write_solution(m, made_solution(), path, mip_relative_gap=0.01)
    h = read_sol(path)["header"]
    for key in ("model", "sense", "status", "algorithm", "objective", "dual_bound",
                "objective_offset", "rows", "columns", "iterations", "primal_infeasibility",
                "dual_infeasibility", "certificate", "source", "nodes", "solve_seconds",
                "mip_relative_gap", "mip_absolute_gap", "integrality_violation"):
        assert key in h, key
    assert h["model"] == "WRITERTEST" and h["source"] == "in-memory"
    assert h["status"] == "optimal" and h["sense"] == "maximize"
    assert h["algorithm"] == "pdhg-gpu" and h["certificate"] == "none"
    assert h["rows"] == "3" and h["columns"] == "3" and h["iterations"] == "1234"
    assert float(h["objective_offset"]) == 2.5
    assert float(h["mip_relative_gap"]) == 0.01 and float(h["mip_absolute_gap"]) == 1e-6


def test_negative_zero_is_never_printed(tmp_path):
    path = tmp_path / "z.sol"
    sh.write_solution(writertest(), made_solution(objective=-0.0, dual_bound=-0.0), path)
    text = path.read_text()
    assert " -0\n" not in text and " -0 " not in text
    assert read_sol(path)["header"]["objective"] == "0"


def test_infinities_are_printed_as_words(tmp_path):
    path = tmp_path / "i.sol"
    sh.write_solution(writertest(), made_solution(status="iteration_limit", dual_bound=-INF,
                                                  objective=INF), path)
    h = read_sol(path)["header"]
    assert h["dual_bound"] == "-inf" and h["objective"] == "inf"
    assert "1e+30" not in path.read_text()


def test_an_unwritable_path_is_reported(tmp_path):
    with pytest.raises(OSError, match="cannot open"):
        sh.write_solution(writertest(), made_solution(), tmp_path / "no_such_dir" / "deeper" / "x.sol")


def test_names_with_whitespace_are_quoted_and_recovered(tmp_path):
    m = writertest(col_names=("AL PRIME", 'Q"T', "MU"), row_names=("THRUPUT 1", "BACK\\SLASH", "SULPHUR"))
    path = tmp_path / "q.sol"
    s = made_solution()
    sh.write_solution(m, s, path)
    text = path.read_text()
    assert '"AL PRIME" ' in text and '"THRUPUT 1" ' in text and '"Q\\"T" ' in text
    f = read_sol(path)
    assert f["column_order"] == ["AL PRIME", 'Q"T', "MU"]
    assert f["row_order"] == ["THRUPUT 1", "BACK\\SLASH", "SULPHUR"]
    assert f["columns"]["AL PRIME"][0] == s.col_value[0]
    assert f["rows"]["THRUPUT 1"][0] == s.row_activity[0]


def test_ordinary_and_default_names_are_not_quoted(tmp_path):
    path = tmp_path / "o.sol"
    sh.write_solution(writertest(), made_solution(), path)
    assert '"' not in path.read_text()
    bare = sh.Model.from_dense([[1.0, 2.0]], [1.0], [1.0], [1.0, 1.0])
    path2 = tmp_path / "d.sol"
    sh.write_solution(bare, made_solution(col_value=np.zeros(2), col_dual=np.zeros(2),
                                          row_dual=np.zeros(1), row_activity=np.zeros(1)), path2)
    f = read_sol(path2)
    assert f["column_order"] == ["C0", "C1"] and f["row_order"] == ["R0"]
    assert f["header"]["model"] == "(unnamed)"


@pytest.mark.parametrize("status", ["numerical_error", "model_error", "not_solved"])
def test_a_status_that_claims_no_point_writes_no_point(tmp_path, status):
    path = tmp_path / "n.sol"
    sh.write_solution(writertest(), made_solution(status=status, message="device lost"), path)
    f = read_sol(path)
    assert f["header"]["status"] == status and f["header"]["message"] == "device lost"
    assert f["declared"] == {} and not f["columns"] and not f["rows"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["afiro", "sc50a"])
def test_engine_solution_file_passes_the_acceptance_check(tmp_path, name):
    cm.gpu_or_skip()
    d = cm.netlib(name)
    mdl = cm.model_of(d)
    mdl.set_names(name.upper(), f"netlib/{name}.mps", d["col_names"], d["row_names"])
    s = sh.solve(mdl, pdhg_tolerance=1e-8, iteration_limit=1000000, pdhg_polish=False)
    assert s.status == "optimal", s.message
    path = tmp_path / f"{name}.sol"
    sh.write_solution(mdl, s, path)
    f = read_sol(path)
    h = f["header"]
    assert h["status"] == "optimal" and h["certificate"] == "none" and h["algorithm"] == s.algorithm
    assert int(h["iterations"]) == s.iterations and float(h["objective"]) == s.objective
    assert f["column_order"] == d["col_names"] and f["row_order"] == d["row_names"]
    from_file = dict(status=h["status"], objective=float(h["objective"]),
                     col_value=np.array([f["columns"][c][0] for c in d["col_names"]]),
                     col_dual=np.array([f["columns"][c][1] for c in d["col_names"]]),
                     row_activity=np.array([f["rows"][r][0] for r in d["row_names"]]),
                     row_dual=np.array([f["rows"][r][1] for r in d["row_names"]]))
    assert np.array_equal(from_file["col_value"], s.col_value)
    assert np.array_equal(from_file["row_dual"], s.row_dual)
    ok, checks = accepted(d, from_file)
    assert ok, [c for c in checks if not c[1]]
    print(f"{name}: {path.name} accepted, objective {h['objective']}", flush=True)
