"""The C++ mirror of the reference's stepping interfaces (mfcm_amd/host/).

CPU tests: the headers compile on their own, the driver is built and links
against libmfsv_hip.so, and its error handling follows the reference's
(usage error, a bad input file, no device: a message and a non-zero exit, no
crash).  The RK tables of the C++ mirror equal the Python front end's.
GPU tests: the driver's whole run equals the CPU oracle to the parity bar of
test_gpu_parity.py, and the Python front end bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from mfcm_amd import drive_io
from mfcm_amd import scenarios as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "mfcm_amd", "host")


def _driver():
    if not os.path.exists(drive_io.DRIVER):
        from mfcm_amd import build as B
        if not os.path.exists(B.OUT):
            pytest.skip("libmfsv_hip.so not built")
        B.build_driver()
    return drive_io.DRIVER


@pytest.mark.parametrize("hdr", ["DataArray.hpp", "TemporalScheme.hpp", "RungeKutta.hpp",
                                 "SaintVenantSolver.hpp"])
def test_headers_compile_alone(hdr, tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(f'#include "{os.path.join(HOST, hdr)}"\nint main() {{ return 0; }}\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_rk_tables_match_python_front_end(tmp_path):
    from mfcm_amd.mfsv import TABLEAUX, tableau_matrix
    names = sorted(TABLEAUX)
    src = tmp_path / "t.cpp"
    body = "\n".join(
        f'  {{ auto m = mfsv::named_coefficients("{n}").matrix(); '
        f'std::printf("{n}"); for (double x : m) std::printf(" %.17g", x); std::printf("\\n"); }}'
        for n in names)
    src.write_text(f'#include <cstdio>\n#include "{os.path.join(HOST, "RungeKutta.hpp")}"\n'
                   f"int main() {{\n{body}\n  return 0;\n}}\n")
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    for line in out.strip().splitlines():
        name, *vals = line.split()
        _, m = tableau_matrix(name)
        np.testing.assert_array_equal(np.array([float(v) for v in vals]), m.ravel())


def test_generic_order3_family_satisfies_order_conditions(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(f'#include <cstdio>\n#include "{os.path.join(HOST, "RungeKutta.hpp")}"\n'
                   "int main() { for (double al : {0.3, 0.5, 0.8}) {"
                   " mfsv::RungeKuttaCoefficientSet s(3, al);"
                   " double b1=s.b(1),b2=s.b(2),b3=s.b(3),c2=s.c(2),c3=s.c(3),a32=s.a(3,2);"
                   " std::printf(\"%.17g %.17g %.17g %.17g\\n\", b1+b2+b3, b2*c2+b3*c3,"
                   " b2*c2*c2+b3*c3*c3, b3*a32*c2); } return 0; }\n")
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    for line in out.strip().splitlines():
        s1, s2, s3, s4 = (float(x) for x in line.split())
        assert s1 == pytest.approx(1.0, abs=1e-14)
        assert s2 == pytest.approx(0.5, abs=1e-14)
        assert s3 == pytest.approx(1.0 / 3.0, abs=1e-14)
        assert s4 == pytest.approx(1.0 / 6.0, abs=1e-14)


def test_driver_usage_and_errors(tmp_path):
    exe = _driver()
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([exe, str(tmp_path / "missing"), str(tmp_path)], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_driver_without_device_fails_cleanly(tmp_path):
    """No GPU here: mfsv_create reports a HIP error, which the mirror turns
    into std::runtime_error and the driver into exit status 1."""
    from mfcm_amd import mfsv
    if mfsv.device_count() > 0:
        pytest.skip("a device is visible")
    exe = _driver()
    sc = S.dam_break_bump(8, 6)
    drive_io.write_input(sc, str(tmp_path / "in"))
    r = subprocess.run([exe, str(tmp_path / "in"), str(tmp_path / "out")], capture_output=True,
                       text=True)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "mfsv_create failed" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["dambreak", "hillslope_prewet_parity", "storm_physics"])
def test_driver_run_matches_oracle(gpu_lib, case, tmp_path):
    from oracle import oracle as O
    if case == "dambreak":
        sc = S.dam_break_bump(64, 48)
        sc.t_end = 60.0
    elif case == "hillslope_prewet_parity":
        sc = S.hillslope_rain(seed=0, n=48, dx=5.0, prewet=0.002, t_end=240.0)
    else:
        sc = S.storm_catchment(seed=1, n=64, dx=2.0, t_end=120.0)
    print(f"[host-cpp] driver {case} {sc.nx}x{sc.ny} t_end={sc.t_end} flags={sc.flags}", flush=True)
    drive_io.write_input(sc, str(tmp_path / "in"))
    r = drive_io.run_driver(str(tmp_path / "in"), str(tmp_path / "out"))
    assert r.returncode == 0, r.stderr
    h, u, v, st = drive_io.read_output(str(tmp_path / "out"), sc.nx, sc.ny)
    sim = O.sim_from_scenario(sc)
    sim.solve()
    ho, uo, vo = sim.get_state(0)
    acc, rej = sim.stats()
    assert (st["timesteps"], st["repeated_timesteps"]) == (acc, rej)
    assert np.abs(h - ho).max() <= 1e-9
    assert np.abs(h * u - ho * uo).max() <= 1e-9
    assert np.abs(h * v - ho * vo).max() <= 1e-9
    np.testing.assert_array_equal(h > 1e-6, ho > 1e-6)
    vo_ = sim.volume(0)
    assert abs(st["volume"] - vo_) <= 1e-10 * max(abs(vo_), 1e-30)


@pytest.mark.gpu
def test_driver_equals_python_front_end_fp32(gpu_lib, tmp_path):
    from mfcm_amd.mfsv import F32, Catchment
    sc = S.hillslope_rain(seed=2, n=40, dx=5.0, physics=True, prewet=0.001, t_end=180.0)
    print("[host-cpp] driver vs Catchment, f32", flush=True)
    drive_io.write_input(sc, str(tmp_path / "in"), precision="f32")
    r = drive_io.run_driver(str(tmp_path / "in"), str(tmp_path / "out"))
    assert r.returncode == 0, r.stderr
    h, u, v, st = drive_io.read_output(str(tmp_path / "out"), sc.nx, sc.ny)
    cat = Catchment(sc, precision=F32)
    cat.solve()
    for a, b in zip((h, u, v), cat.state()):
        np.testing.assert_array_equal(a, b)
    assert (st["timesteps"], st["repeated_timesteps"]) == (cat.n_acc, cat.n_rej)
    cat.close()


_CFG = """mesh == cartesian
{
  cell count == 8, 6
  cell size == 2.0, 2.0
}
scheme == runge-kutta
{
  time unit == s
  end time == 60
}
"""


@pytest.mark.parametrize("extra,msg", [
    ("source term == magic\n", "Unknown source term type"),
    ("source term == manning roughness\nsource term == manning roughness\n", "configured twice"),
    ("boundary == head\nboundary == discharge\n", "order discharge, head, stage"),
    ("boundary == discharge\nboundary == discharge\n", "more than one"),
    ("boundary == flood\n", "Unknown boundary type"),
    ("boundary == discharge\n{\n  time series == r\n  {\n    values == nope\n    cells == global\n"
     "  }\n}\n", "Could not find time series 'nope'"),
    ("output == o\n{\n  at == edges\n  every seconds == 1\n}\n", "output.at"),
])
def test_driver_config_mode_reports_configuration_errors(tmp_path, extra, msg):
    """Configuration mode: errors in the configuration end the run with a
    message and status 1 before any device work."""
    exe = _driver()
    cfg = tmp_path / "a.mf"
    cfg.write_text(_CFG + extra)
    r = subprocess.run([exe, "--quiet", str(cfg)], capture_output=True, text=True)
    assert r.returncode == 1 and msg in r.stderr, (r.stdout, r.stderr)


def test_driver_config_mode_usage(tmp_path):
    exe = _driver()
    r = subprocess.run([exe, "--bogus", "x.mf"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = subprocess.run([exe, "--quiet", str(tmp_path / "none.mf")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open file" in r.stderr
