"""The input and output layer of the C++ mirror (mfcm_amd/host/Config.hpp,
Input.hpp, Output.hpp) on the CPU: the reference's mf configuration format,
times, time series, mesh selections, field generators, NIMROD rasters, mesh
locations and the CSV output format.

The reference ships no tests or fixtures for these, so each case states its
expectation from the reference's source text: the grammar of
Config/mf_parser/mf_parser_read.hpp, the unit table and defaults of
Config/TimeParser.cpp and TimeParameters.cpp, the selection and generator
rules of Mesh/MeshSelection.cpp and Input/FieldGenerator.cpp, the transform of
Mesh/Cartesian2DMesh.cpp, the NIMROD record layout of
Raster/RasterFormats/NIMROD.cpp and the formatting of Output/OutputFormat.hpp.
The probe program tests/host/config_probe.cpp prints what the headers make of
a configuration.
"""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE_SRC = os.path.join(ROOT, "tests", "host", "config_probe.cpp")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("probe") / "config_probe"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(ROOT, "include"), "-o", str(exe), PROBE_SRC], check=True)

    def run(*args, ok=True):
        r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True)
        if ok:
            assert r.returncode == 0, r.stdout + r.stderr
        return r.stdout
    return run


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(text)
    return p


MESH = """mesh == cartesian
{
  cell count == 4, 3
  cell size == 2.0, 2.0
}
"""


def _tree(out):
    return [tuple(line.split("|", 2)) for line in out.strip().splitlines()]


def test_mf_grammar(probe, tmp_path):
    _write(tmp_path, "sub/inc.mf", "included == yes\n")
    f = _write(tmp_path, "main.mf", r"""! a comment line
mesh == cartesian   ! trailing comment
{
  cell count == 4, 3
  cell size == 2.0,  2.0
}
multi word   key == some value with  spaces
"quoted key" == v
text == "a \"quoted\" string" \
        " continued"
bare key words
block ==
{
  inner == 1
  { nested-open-after-entry
  }
}
esc == tab\there
glued {
#include "sub/inc.mf"
""")
    t = _tree(probe("tree", f))
    assert t == [
        ("0", "mesh", "cartesian"),
        ("1", "cell count", "4, 3"),
        ("1", "cell size", "2.0,  2.0"),
        ("0", "multi word key", "some value with  spaces"),
        # a quoted key leaves the delimiter in the value (read_key/read_data)
        ("0", "quoted key", "== v"),
        ("0", "text", 'a "quoted" string continued'),
        ("0", "bare key words", ""),
        ("0", "block", ""),
        ("1", "inner", "1"),
        ("2", "nested-open-after-entry", ""),
        ("0", "esc", "tab\there"),
        # a brace after a key on the same line without "==" is key text
        ("0", "glued {", ""),
        ("0", "included", "yes"),
    ]


@pytest.mark.parametrize("text,msg", [
    ("a == 1\n}\n", "unmatched }"),
    ("a == 1\n{\n", "unmatched {"),
    ("{\n", "unexpected {"),
    ("#define x\n", "unknown directive"),
    ('#include "nope.mf"\n', "cannot open include file nope.mf"),
    ('a == "open\n', "unexpected end of line"),
    ("a == bad\\q\n", "unknown escape sequence"),
    ('a == "x" \\\nb == 2\n', "expected \" after \\ in previous line"),
])
def test_mf_errors_name_file_and_line(probe, tmp_path, text, msg):
    f = _write(tmp_path, "bad.mf", text)
    out = probe("tree", f, ok=False)
    assert out.startswith("error: ") and msg in out, out
    assert str(f) + "(" in out


def test_time_parser_units_and_formats(probe, tmp_path):
    base = MESH + "scheme == runge-kutta\n{\n  end time == 2\n%s}\n"
    f = _write(tmp_path, "a.mf", base % "")
    assert float(probe("time", f, "1.5")) == 5400.0          # hours unless a unit is given
    f = _write(tmp_path, "b.mf", base % "  time unit == min\n")
    assert float(probe("time", f, "1.5")) == 90.0
    f = _write(tmp_path, "c.mf", base % "  time unit == 0.5\n")  # a factor to seconds
    assert float(probe("time", f, "3")) == 1.5
    f = _write(tmp_path, "d.mf", base % "  time unit == fortnights\n")
    assert "Unknown time unit" in probe("time", f, "1", ok=False)
    f = _write(tmp_path, "e.mf", base % "  time format == %Y-%m-%d %H:%M:%S\n"
                                        "  time zero == 2020-01-01 00:00:00\n")
    assert float(probe("time", f, "2020-01-01 02:30:00")) == 9000.0
    # a number is still a number (in the unit) when it does not match the format
    assert float(probe("time", f, "0.5")) == 1800.0


def test_time_parameters_defaults_and_step_count(probe, tmp_path):
    f = _write(tmp_path, "a.mf", MESH + "scheme == runge-kutta\n{\n  time unit == s\n"
                                        "  end time == 5400\n  step duration seconds == 700\n}\n")
    start, end, step, dt0, n = probe("tparams", f).split()
    assert (float(start), float(end), float(step), float(dt0)) == (0.0, 5400.0, 700.0, 1.0)
    assert int(n) == 8                     # 5400 / 700 = 7.7 rounds up
    f = _write(tmp_path, "b.mf", MESH + "scheme == runge-kutta\n{\n  time unit == s\n"
                                        "  start time == 100\n  end time == 400\n"
                                        "  step duration == 100\n  initial timestep seconds == 0.25\n}\n")
    start, end, step, dt0, n = probe("tparams", f).split()
    assert (float(start), float(end), float(step), float(dt0), int(n)) == (100.0, 400.0, 100.0, 0.25, 3)
    f = _write(tmp_path, "c.mf", MESH + "scheme == runge-kutta\n{\n  end time == 1\n}\n")
    assert probe("tparams", f).split()[2] == "60"  # default step duration: 60 s


SERIES_HEAD = MESH + "scheme == runge-kutta\n{\n  time unit == s\n  end time == 100\n}\n"


def test_inline_series_with_transform(probe, tmp_path):
    f = _write(tmp_path, "a.mf", SERIES_HEAD + """time series == rain
{
  source == inline
  {
    time unit == min
    data ==
    {
      0 == 1.0
      1.5 == 3.0
      10 == -2
    }
  }
  transform ==
  {
    time offset == 1
    time factor == 2
    value offset == 0.5
    value factor == 1e-3
  }
}
""")
    rows = [tuple(map(float, l.split())) for l in probe("series", f, "rain").strip().splitlines()]
    assert rows == [((0 + 1) * 2, 1.5e-3), ((90 + 1) * 2, 3.5e-3), ((600 + 1) * 2, -1.5e-3)]


@pytest.mark.parametrize("data,msg", [
    ("      0 == 1\n      0 == 2\n", "must increase"),
    ("      5 == 1\n", "at least two"),
])
def test_inline_series_errors(probe, tmp_path, data, msg):
    f = _write(tmp_path, "a.mf", SERIES_HEAD + "time series == s\n{\n  source == inline\n  {\n"
               "    data ==\n    {\n" + data + "    }\n  }\n}\n")
    assert msg in probe("series", f, "s", ok=False)


def test_csv_series_headers_comments_quotes(probe, tmp_path):
    _write(tmp_path, "rain.csv", "# gauge export\nheader junk\n"
           ' time ;"depth ; mm" ; other\n0;1.5;x\n# a comment row\n\n60 ; "2.5";y\n120;bad;z')
    f = _write(tmp_path, "a.mf", SERIES_HEAD + """time series == g
{
  source == csv
  filename == rain.csv
  separator == ;
  skip rows == 0
  time column == time
  value column == depth ; mm
}
""")
    # the comment line is dropped before rows are counted: "header junk" is
    # row 0 (skipped), the names are row 1; a quote opens a quoted cell only
    # as the cell's first character; a bad number reads as NaN
    rows = [l.split() for l in probe("series", f, "g").strip().splitlines()]
    assert [float(r[0]) for r in rows] == [0.0, 60.0, 120.0]
    assert float(rows[0][1]) == 1.5 and float(rows[1][1]) == 2.5 and math.isnan(float(rows[2][1]))


def test_csv_series_without_headers(probe, tmp_path):
    _write(tmp_path, "s.csv", "a,0,10\nb,30,20\nc,60,40\n")
    f = _write(tmp_path, "a.mf", SERIES_HEAD + """time series == s
{
  source == csv
  filename == s.csv
  headers == false
  time column == 2
  value column == 3
}
""")
    rows = [tuple(map(float, l.split())) for l in probe("series", f, "s").strip().splitlines()]
    assert rows == [(0.0, 10.0), (30.0, 20.0), (60.0, 40.0)]


SELECT = MESH + "sel == %s\n"


@pytest.mark.parametrize("body,expect", [
    ("", "global"),
    ("global", "global"),
    # id entries and id lists, sorted, duplicates and ids past the mesh dropped
    ("id list\n{\n  id == 7\n  ids == 3,1,3\n  id == 100\n}", "ids 1 3 7"),
    # the cell containing each point: (5, 3) lies in column 2 of row 1 (id
    # 6), (0.5, 5.9) in column 0 of row 2 (id 8); (100, 100) is outside
    ("at\n{\n  wkt == POINT (5 3), point(0.5 5.9), POINT (100 100)\n}", "ids 6 8"),
])
def test_selections(probe, tmp_path, body, expect):
    text = SELECT % body
    f = _write(tmp_path, "a.mf", text)
    assert probe("select", f, "sel").strip() == expect


@pytest.mark.parametrize("body,msg", [
    ("id list\n{\n  ids == 1, 2\n}", "bad lexical cast"),   # numbers as written, no blanks
    ("at\n{\n  wkt == MULTIPOINT ((1 1))\n}", "Cannot update mask"),
    ("nearby", "Unknown mesh selection method"),
])
def test_selection_errors(probe, tmp_path, body, msg):
    f = _write(tmp_path, "a.mf", SELECT % body)
    assert msg in probe("select", f, "sel", ok=False)


def _mesh_xy(nx, ny, ox=0.0, oy=0.0, sx=2.0, sy=2.0, ang=0.0):
    b, c = sx * math.cos(ang), sx * math.sin(ang)
    e, f = sy * math.sin(ang), sy * math.cos(ang)
    def loc(xi, yi):
        return ox + xi * b + yi * c, oy + xi * e + yi * f
    return loc


def test_field_generator_operations(probe, tmp_path):
    f = _write(tmp_path, "a.mf", MESH + """field == z_bed
{
  set == slope
  {
    origin == 1.0, 2.0, 10.0
    slope == 0.1, -0.25
  }
  add == 1.5
  {
    within == id list
    {
      ids == 0,5
    }
  }
  multiply == constant
  {
    value == 2
  }
  add == composite
  {
    model == constant
    {
      value == 7
      cells == id list
      {
        id == 11
      }
    }
  }
}
""")
    got = np.array([float(x) for x in probe("field", f, "z_bed", "0").split()])
    loc = _mesh_xy(4, 3)
    exp = np.zeros(12)
    for i in range(12):
        x, y = loc(i % 4 + 0.5, i // 4 + 0.5)
        exp[i] = 10.0 + 0.1 * (x - 1.0) + -0.25 * (y - 2.0)
    exp[[0, 5]] += 1.5
    exp *= 2.0
    exp[11] += 7.0
    np.testing.assert_array_equal(got, exp)
    # a field without a section keeps its default
    assert [float(x) for x in probe("field", f, "h", "0.25").split()] == [0.25] * 12


def test_field_generator_unknown_specifier(probe, tmp_path):
    f = _write(tmp_path, "a.mf", MESH + "field == h\n{\n  set == lots\n}\n")
    assert "not a known input field specifier" in probe("field", f, "h", "0", ok=False)


def test_mesh_transform_and_locations(probe, tmp_path):
    ang = 0.3
    f = _write(tmp_path, "a.mf", "mesh == cartesian\n{\n  cell count == 3, 2\n"
               f"  origin == 10.0, -5.0\n  cell size == 2.0, 1.5\n  angle == {ang!r}\n}}\n")
    lines = probe("mesh", f).strip().splitlines()
    dx, dy = map(float, lines[0].split())
    b, c = 2.0 * math.cos(ang), 2.0 * math.sin(ang)
    e, fy = 1.5 * math.sin(ang), 1.5 * math.cos(ang)
    assert dx == math.sqrt(b * b + c * c) and dy == math.sqrt(e * e + fy * fy)
    loc = _mesh_xy(3, 2, 10.0, -5.0, 2.0, 1.5, ang)
    want = [("c",) + loc(i % 3 + 0.5, i // 3 + 0.5) for i in range(6)]
    want += [("f",) + loc(i % 4, i // 4 + 0.5) for i in range(8)]       # x-faces
    want += [("f",) + loc(i % 3 + 0.5, i // 3) for i in range(9)]       # y-faces
    want += [("v",) + loc(i % 4, i // 4) for i in range(12)]
    got = [(l.split()[0], float(l.split()[1]), float(l.split()[2])) for l in lines[1:]]
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g[0] == w[0] and g[1] == pytest.approx(w[1], abs=1e-12) and g[2] == pytest.approx(w[2], abs=1e-12)


def _nimrod(path, data, ncols, nrows, x_first, y_top, px, py, dtype="h"):
    """A NIMROD file: 4-byte big-endian record lengths around a 512-byte
    header (31 int16, 28 float32, 45 float32, 56 chars, 51 int16) and the
    data record, row 0 the northernmost."""
    h1 = [0] * 31
    h1[11] = 0 if dtype == "f" else 1
    h1[12] = 4 if dtype == "f" else 2
    h1[14] = 0                        # National Grid
    h1[15], h1[16] = nrows, ncols
    h1[23] = 0                        # top-left origin
    h1[24] = -1
    h2 = [0.0] * 28
    h2[2], h2[3], h2[4], h2[5], h2[6] = y_top, py, x_first, px, -1.0
    head = (struct.pack(">31h", *h1) + struct.pack(">28f", *h2) + struct.pack(">45f", *([0.0] * 45))
            + b" " * 56 + struct.pack(">51h", *([0] * 51)))
    assert len(head) == 512
    body = struct.pack(f">{len(data)}{dtype}", *data)
    with open(path, "wb") as fh:
        fh.write(struct.pack(">I", 512) + head + struct.pack(">I", 512))
        fh.write(struct.pack(">I", len(body)) + body + struct.pack(">I", len(body)))


@pytest.mark.parametrize("dtype,bbox", [("h", None), ("f", None), ("h", "5.0, 1.0, 14.0, 9.0")])
def test_nimrod_raster_field(probe, tmp_path, dtype, bbox):
    # 5 x 4 pixels of 4 m; first column centred at x = 2, top row at y = 14
    ncols, nrows, px = 5, 4, 4.0
    vals = [10 * r + c for r in range(nrows) for c in range(ncols)]
    _nimrod(tmp_path / "radar.dat", vals, ncols, nrows, 2.0, 14.0, px, px, dtype)
    extra = f"  bbox == {bbox}\n" if bbox else ""
    f = _write(tmp_path, "a.mf", "mesh == cartesian\n{\n  cell count == 6, 4\n"
               "  origin == 1.0, 0.5\n  cell size == 3.0, 3.0\n}\n"
               "raster field == radar\n{\n  source == NIMROD\n  filename == radar.dat\n" + extra + "}\n"
               "field == IR\n{\n  set == raster\n  {\n    name == radar\n  }\n}\n")
    got = [float(x) for x in probe("field", f, "IR", "-9").split()]
    x_left, y_bottom = 2.0 - px / 2, 14.0 + px / 2 - nrows * px
    c0, c1, r0, r1 = 0, ncols - 1, 0, nrows - 1
    if bbox:
        bx0, by0, bx1, by1 = map(float, bbox.split(","))
        c0, c1 = int((bx0 - x_left) / px), int((bx1 - x_left) / px)
        r0, r1 = int((16.0 - by1) / px), int((16.0 - by0) / px)
    for i, g in enumerate(got):
        x = 1.0 + (i % 6 + 0.5) * 3.0
        y = 0.5 + (i // 6 + 0.5) * 3.0
        col, row = int((x - x_left) // px), int((y_bottom + nrows * px - y) // px)
        inside = (x >= x_left + c0 * px and x < x_left + (c1 + 1) * px and
                  y < 16.0 - r0 * px and y >= 16.0 - (r1 + 1) * px)
        assert g == (10 * row + col if inside else -9.0), (i, x, y, g)


def test_nimrod_bad_header(probe, tmp_path):
    (tmp_path / "radar.dat").write_bytes(struct.pack(">I", 256) + b"\0" * 600)
    f = _write(tmp_path, "a.mf", MESH + "raster field == r\n{\n  source == nimrod\n"
               "  filename == radar.dat\n}\nfield == h\n{\n  set == raster\n  {\n    name == r\n  }\n}\n")
    assert "512" in probe("field", f, "h", "0", ok=False)


def test_csv_format(probe, tmp_path):
    out = tmp_path / "t.csv"
    probe("csv", out)
    # header: "X", "Y" and quoted names, each followed by the delimiter;
    # locations with the stream's default 6 significant digits, values with
    # std::to_string's six decimals
    assert out.read_text() == ('X, Y, "h", "u v", \n'
                               '101.25, 201.25, 1.500000, 0.000000, \n'
                               '103.75, 201.25, -2.250000, 12345.678901, \n')


def test_timed_output_schedule_and_columns(probe, tmp_path):
    f = _write(tmp_path, "a.mf", MESH + """scheme == runge-kutta
{
  time unit == s
  end time == 100
}
output == depth
{
  every == 30
  start time == 10
  field == u
  field == none
  field == h
}
output == faces
{
  every seconds == 1
  at == faces
  field == h
}
output == off
{
  enabled == false
  every seconds == 1
}
""")
    out = probe("output", f, "depth", 0, 10, 20, 40, 45, 70, 100).split()
    names = [os.path.basename(p) for p in out]
    # due when t >= start + every * (outputs so far)
    assert names == ["depth_10.000000.csv", "depth_40.000000.csv", "depth_70.000000.csv",
                     "depth_100.000000.csv"]
    text = (tmp_path / "output" / "depth_10.000000.csv").read_text().splitlines()
    assert text[0] == 'X, Y, "u", "h", '           # a field the solver lacks is left out
    assert text[1] == "1, 1, 0.500000, 0.000000, "
    assert len(text) == 13
    faces = probe("output", f, "faces", 0).split()
    ft = open(faces[0]).read().splitlines()
    assert len(ft) == 1 + 5 * 3 + 4 * 4 and ft[1] == "0, 1, 0.000000, "
    assert probe("output", f, "off", 0, 1, 2).strip() == ""


def test_default_check_file_sections(probe, tmp_path):
    f = _write(tmp_path, "a.mf", MESH)
    assert probe("checkdefaults", f).split() == ["true", "check/constants.csv", "false"]
