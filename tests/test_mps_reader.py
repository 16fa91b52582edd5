"""The library's MPS reader (sankhya::io::read_mps, src/io/mps_reader.cpp and
src/io/mps_bounds.cpp in the reference) through the C API sankhya_model_read (reference
src/api/c_api.cpp:185-200).

Pinned three ways: on the nine committed Netlib files it gives the same arrays, bit for bit, as
the tests' own reader (tests/mps.py, itself pinned against the published Netlib dimensions and
optima); the model it reads has the published Netlib optimum under HiGHS's dual simplex; and
small hand-written files check each rule of the format (RANGES by row type and sign, bound types,
markers, objective sense and constant, vector names, comments, both dialects) against values
worked out by hand from IBM's MPS description, plus every error the reference reports.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np
import pytest

import common as cm
import sankhya_hip as sh
from mps import read_mps

NETLIB = sorted(f[:-4] for f in os.listdir(cm.NETLIB) if f.endswith(".mps"))
INF = math.inf


def _write(tmp_path, text, name="m.mps"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _read(path):
    return sh.Model.read_mps(path)


@pytest.mark.parametrize("name", NETLIB)
def test_netlib_arrays_equal_the_test_reader_bit_for_bit(name):
    path = os.path.join(cm.NETLIB, f"{name}.mps")
    want = read_mps(path)
    model = _read(path)
    got = model.arrays()
    assert (got["m"], got["n"]) == (want["m"], want["n"])
    for key in ("colptr", "rowidx", "val", "cost", "col_lower", "col_upper", "row_lower", "row_upper"):
        a, b = np.asarray(got[key]), np.asarray(want[key])
        assert a.shape == b.shape and a.tobytes() == b.tobytes(), (name, key)
    assert got["objective_offset"] == want["objective_offset"] and not got["maximize"]
    names = model.names()
    assert names["name"] == want["name"] and names["source_path"] == path
    assert names["col_names"] == want["col_names"] and names["row_names"] == want["row_names"]
    assert model.integer_columns().size == 0


@pytest.mark.parametrize("name", NETLIB)
def test_netlib_read_has_the_published_optimum(name):
    d = _read(os.path.join(cm.NETLIB, f"{name}.mps")).arrays()
    pub = cm.netlib_optima()[name]["published_optimal"]
    assert abs(cm.simplex_objective(d) - pub) / max(1.0, abs(pub)) < 1e-7


RANGES = """NAME          RNG
ROWS
 N  obj
 E  e_pos
 E  e_neg
 E  e_zero
 L  l_row
 G  g_row
 L  l_plain
 G  g_plain
 E  e_plain
COLUMNS
    x  obj  1.0  e_pos  1.0
    x  e_neg  1.0  e_zero  1.0
    x  l_row  1.0  g_row  1.0
    x  l_plain  1.0  g_plain  1.0
    x  e_plain  1.0
RHS
    RHS  e_pos  4.0  e_neg  4.0
    RHS  e_zero  4.0  l_row  4.0
    RHS  g_row  4.0  l_plain  4.0
    RHS  g_plain  4.0  e_plain  4.0
RANGES
    RNG  e_pos  3.0  e_neg  -3.0
    RNG  e_zero  0.0  l_row  -3.0
    RNG  g_row  -3.0
ENDATA
"""


def test_ranges_follow_the_ibm_table(tmp_path):
    d = _read(_write(tmp_path, RANGES)).arrays()
    # E, R > 0: [b, b+R]; E, R < 0: [b+R, b]; E, R = 0: [b, b]; L: [b-|R|, b]; G: [b, b+|R|];
    # without RANGES: L (-inf, b], G [b, inf), E [b, b].
    assert list(d["row_lower"]) == [4.0, 1.0, 4.0, 1.0, 4.0, -INF, 4.0, 4.0]
    assert list(d["row_upper"]) == [7.0, 4.0, 4.0, 4.0, 7.0, 4.0, INF, 4.0]


BOUNDS = """NAME BND
ROWS
 N obj
 L c1
COLUMNS
 up_neg obj 1 c1 1
 up_neg_lo obj 1 c1 1
 up_pos obj 1 c1 1
 mi obj 1 c1 1
 fr obj 1 c1 1
 fx obj 1 c1 1
 pl obj 1 c1 1
 bv obj 1 c1 1
 li obj 1 c1 1
 ui_neg obj 1 c1 1
 big obj 1 c1 1
 MARKER 'MARKER' 'INTORG'
 intcol obj 1 c1 1
 MARKER 'MARKER' 'INTEND'
 after obj 1 c1 1
RHS
 RHS c1 10
BOUNDS
 LO BND up_neg_lo -2
 UP BND up_neg -5
 UP BND up_neg_lo -1
 UP BND up_pos 3
 MI BND mi
 UP BND mi 6
 FR BND fr
 FX BND fx 2.5
 UP BND pl 9
 PL BND pl
 BV BND bv
 LI BND li 2
 UI BND ui_neg -4
 LO BND big -1e30
 UP BND big 1e31
 UP OTHER up_pos 100
ENDATA
"""


def test_bound_types(tmp_path):
    model = _read(_write(tmp_path, BOUNDS))
    d = model.arrays()
    lo, up = list(d["col_lower"]), list(d["col_upper"])
    # A negative UP (or UI) on a column with no explicit lower bound frees it below; with an
    # explicit LO it does not. MI leaves the upper bound; later vectors (OTHER) are ignored.
    assert lo == [-INF, -2.0, 0.0, -INF, -INF, 2.5, 0.0, 0.0, 2.0, -INF, -INF, 0.0, 0.0]
    assert up == [-5.0, -1.0, 3.0, 6.0, INF, 2.5, INF, 1.0, INF, -4.0, INF, INF, INF]
    names = model.names()["col_names"]
    assert [names[j] for j in model.integer_columns()] == ["bv", "li", "ui_neg", "intcol"]


def test_objective_sense_constant_and_free_rows(tmp_path):
    text = """* a comment line
NAME          SENSE
OBJSENSE
    MAX
ROWS
 N  profit
 N  spare
 G  c1
COLUMNS
    x  profit  2.0  spare  7.0
    x  c1  1.0D0
    y  profit  3.0  c1  1.0  $ a free-format comment
RHS
    RHS  profit  -12.5  c1  1.0
    OTHER  c1  99.0
ENDATA
"""
    d = _read(_write(tmp_path, text)).arrays()
    assert d["maximize"] and d["objective_offset"] == 12.5  # RHS on the objective is -constant
    assert (d["m"], d["n"]) == (1, 2) and list(d["val"]) == [1.0, 1.0]  # the spare N row dropped
    assert list(d["cost"]) == [2.0, 3.0] and list(d["row_lower"]) == [1.0]
    inline = text.replace("OBJSENSE\n    MAX\n", "OBJSENSE    MAXIMIZE\n")
    assert _read(_write(tmp_path, inline, "inline.mps")).arrays()["maximize"]


def test_zero_drop_and_rows_ascend(tmp_path):
    # Entries under the zero-drop tolerance (1e-11, include/sankhya/tolerances.hpp:126) are no
    # entries, explicit zeros included; 1e-10 is kept.
    text = """NAME Z
ROWS
 N obj
 E r0
 E r1
 E r2
COLUMNS
 x r2 3 r0 1
 x r1 0
 y r1 2 r0 1e-12
 y r2 1e-10
 x obj 5
ENDATA
"""
    d = _read(_write(tmp_path, text)).arrays()
    assert list(d["colptr"]) == [0, 2, 4] and list(d["rowidx"]) == [0, 2, 1, 2]
    assert list(d["val"]) == [1.0, 3.0, 2.0, 1e-10] and list(d["cost"]) == [5.0, 0.0]


def _fixed(*fields):
    """One fixed-format data line: fields at 1-based columns 2, 5, 15, 25, 40 and 50."""
    line = ""
    for at, f in zip((1, 4, 14, 24, 39, 49), fields):
        line = line.ljust(at) + f
    return line + "\n"


FIXED = ("NAME          FIX\nROWS\n" + _fixed("N", "COST") + _fixed("L", "LIM 1") + _fixed("G", "LIM 2")
         + "COLUMNS\n" + _fixed("", "X ONE", "COST", "1.0", "LIM 1", "1.0")
         + _fixed("", "X ONE", "LIM 2", "1.0") + _fixed("", "X TWO", "COST", "2.0", "LIM 1", "1.0")
         + "RHS\n" + _fixed("", "RHS", "LIM 1", "4.0", "LIM 2", "1.0")
         + "BOUNDS\n" + _fixed("UP", "BND", "X TWO", "3.0") + "ENDATA\n")


def test_fixed_format_names_with_blanks(tmp_path):
    path = _write(tmp_path, FIXED)
    model = _read(path)  # free reading fails on the blanks; the fixed columns read it
    d = model.arrays()
    assert model.names()["col_names"] == ["X ONE", "X TWO"]
    assert model.names()["row_names"] == ["LIM 1", "LIM 2"]
    assert list(d["colptr"]) == [0, 2, 3] and list(d["rowidx"]) == [0, 1, 0]
    assert list(d["row_upper"]) == [4.0, INF] and list(d["row_lower"]) == [-INF, 1.0]
    assert list(d["col_upper"]) == [INF, 3.0] and list(d["cost"]) == [1.0, 2.0]


BASE = """NAME E
ROWS
 N obj
 L c1
COLUMNS
 x obj 1 c1 1
{extra_columns}RHS
 RHS c1 1
{extra}ENDATA
"""


@pytest.mark.parametrize("extra_columns,extra,message", [
    (" x c1 2\n", "", "duplicate entry for row 'c1' and column 'x'"),
    (" x obj 2\n", "", "duplicate objective coefficient for column 'x'"),
    (" x c9 2\n", "", "COLUMNS entry names unknown row 'c9'"),
    (" y c1 1e400\n", "", "must be finite"),
    (" y c1 nan\n", "", "'nan' is not a number"),
    (" y c1 1 c1\n", "", "COLUMNS entry has 4 fields"),
    (" MARKER 'MARKER' 'INTX'\n", "", "MARKER record names neither INTORG nor INTEND"),
    ("", "BOUNDS\n SC BND x 4\n", "semi-continuous bounds (SC) are not supported"),
    ("", "BOUNDS\n XX BND x 4\n", "unknown bound type 'XX'"),
    ("", "BOUNDS\n UP BND z 4\n", "names column 'z', which has no COLUMNS entry"),
    ("", "RANGES\n RNG obj 4\n", "RANGES entry on the objective row is not meaningful"),
    ("", "QUADOBJ\n x x 1\n", "reads linear programs only"),
])
def test_errors_name_the_file_line_and_cause(tmp_path, extra_columns, extra, message):
    path = _write(tmp_path, BASE.format(extra_columns=extra_columns, extra=extra))
    with pytest.raises(ValueError) as e:
        _read(path)
    assert message in str(e.value)
    assert str(e.value).startswith(path + ":")


def test_structural_errors(tmp_path):
    cases = {
        "no_end.mps": BASE.format(extra_columns="", extra="").replace("ENDATA\n", ""),
        # b = 1e30 reads as +inf and R = -1e400 overflows to -inf: the E row's lower end b + R
        # is NaN, which validation refuses.
        "nan_range.mps": BASE.format(extra_columns="", extra="").replace(" L c1", " E c1")
        .replace("ENDATA", "RANGES\n RNG c1 -1e400\nENDATA").replace("RHS c1 1", "RHS c1 1e30"),
        "row_shape.mps": BASE.format(extra_columns="", extra="").replace(" L c1", " L c 1"),
        "data_first.mps": " x obj 1\n" + BASE.format(extra_columns="", extra=""),
        "crossed.mps": BASE.format(extra_columns="", extra="BOUNDS\n LO BND x 5\n UP BND x 4\n"),
    }
    expect = {
        "no_end.mps": "the file ends without an ENDATA record",
        "nan_range.mps": "row 0 has a NaN bound",
        "row_shape.mps": "ROWS entry has 3 fields",
        "data_first.mps": "data line appears before any recognised section header",
        "crossed.mps": "model failed validation",
    }
    for name, text in cases.items():
        with pytest.raises(ValueError) as e:
            _read(_write(tmp_path, text, name))
        assert expect[name] in str(e.value), (name, str(e.value))


def _c_api():
    L = sh.load_library()
    L.sankhya_model_create.restype = C.c_void_p
    L.sankhya_model_read.argtypes = [C.c_void_p, C.c_char_p]
    L.sankhya_model_free.argtypes = [C.c_void_p]
    L.sankhya_model_num_rows.argtypes = [C.c_void_p]
    L.sankhya_model_num_cols.argtypes = [C.c_void_p]
    L.sankhya_last_error.restype = C.c_char_p
    return L


def test_c_api_read_keeps_the_handle_on_failure(tmp_path):
    L = _c_api()
    h = L.sankhya_model_create()
    try:
        assert L.sankhya_model_read(h, os.path.join(cm.NETLIB, "afiro.mps").encode()) == 0
        assert (L.sankhya_model_num_rows(h), L.sankhya_model_num_cols(h)) == (27, 32)
        bad = _write(tmp_path, BASE.format(extra_columns=" x c1 2\n", extra=""))
        assert L.sankhya_model_read(h, bad.encode()) == 2  # SANKHYA_ERROR_IO
        assert b"duplicate entry" in L.sankhya_last_error()
        assert (L.sankhya_model_num_rows(h), L.sankhya_model_num_cols(h)) == (27, 32)
        assert L.sankhya_model_read(h, str(tmp_path / "absent.mps").encode()) == 2
        assert b"cannot open" in L.sankhya_last_error()
        assert L.sankhya_model_read(None, b"x.mps") == 1  # SANKHYA_ERROR_ARGUMENT
    finally:
        L.sankhya_model_free(h)


def test_lp_input_is_refused_by_name(tmp_path):
    for name in ("m.lp", "m.lp.gz"):
        with pytest.raises(ValueError) as e:
            _read(_write(tmp_path, "NAME X\n", name))
        assert "CPLEX LP files are not read" in str(e.value)


@pytest.mark.parametrize("name", ["afiro", "sc50b", "adlittle"])
def test_gzip_input_reads_as_the_plain_file(tmp_path, name):
    """A gzip-compressed MPS file (the reference reads `.gz` through zlib,
    src/io/line_reader.cpp:43-60) gives the plain file's model bit for bit; compressed data cut
    short is an error naming the file, not a shorter model."""
    import gzip

    src = os.path.join(cm.NETLIB, f"{name}.mps")
    raw = open(src, "rb").read()
    gz = tmp_path / f"{name}.mps.gz"
    gz.write_bytes(gzip.compress(raw))
    plain, packed = _read(src).arrays(), _read(str(gz)).arrays()
    assert plain.keys() == packed.keys()
    for key in plain:
        np.testing.assert_array_equal(np.asarray(plain[key]), np.asarray(packed[key]), err_msg=key)
    cut = tmp_path / f"{name}_cut.mps.gz"
    cut.write_bytes(gzip.compress(raw)[: len(gzip.compress(raw)) // 2])
    with pytest.raises(ValueError) as e:
        _read(str(cut))
    assert f"{name}_cut.mps.gz" in str(e.value)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["afiro", "sc50b", "adlittle"])
def test_library_read_solves_on_the_gpu_and_its_solution_file_is_accepted(tmp_path, name):
    """The reference CLI's path end to end: the library reads the MPS file (names included), the
    MI355X engine solves it to 1e-8, the .sol file is written, read back, and accepted by the
    verifier restatement against the tests' own reading of the same file; the objective matches
    the published Netlib optimum."""
    from solfile import read_sol
    from verify import accepted

    cm.gpu_or_skip()
    path = os.path.join(cm.NETLIB, f"{name}.mps")
    model = _read(path)
    s = sh.solve(model, pdhg_tolerance=1e-8, iteration_limit=1_000_000, pdhg_polish=False)
    assert s.status == "optimal", s.message
    pub = cm.netlib_optima()[name]["published_optimal"]
    assert abs(s.objective - pub) / max(1.0, abs(pub)) < 1e-6
    sol = tmp_path / f"{name}.sol"
    sh.write_solution(model, s, sol)
    f = read_sol(sol)
    want = read_mps(path)
    assert f["column_order"] == want["col_names"] and f["row_order"] == want["row_names"]
    point = dict(status=f["header"]["status"], objective=float(f["header"]["objective"]),
                 col_value=np.array([f["columns"][c][0] for c in want["col_names"]]),
                 col_dual=np.array([f["columns"][c][1] for c in want["col_names"]]),
                 row_activity=np.array([f["rows"][r][0] for r in want["row_names"]]),
                 row_dual=np.array([f["rows"][r][1] for r in want["row_names"]]))
    ok, checks = accepted(want, point)
    assert ok, [c for c in checks if not c[1]]
    print(f"{name}: read by the library, {s.iterations} iterations, objective {s.objective:.10g} "
          f"(published {pub}), {sol.name} accepted", flush=True)
