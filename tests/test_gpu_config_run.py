"""A run driven by a configuration file in the reference's own format, through
the C++ driver (mfcm_amd/host/mfsv_drive, config mode) on the GPU, against
the CPU oracle set up with the same bed, state, roughness and boundaries.

Bar (FP64 plain method), as for the whole-run parity tests: final
max |dh|, |d(hu)|, |d(hv)| <= 1e-9, wet masks (h > 1e-6) identical, the same
accepted / repeated timestep counts.  The CSV outputs hold six decimals
(std::to_string), so they are compared to the oracle within 5e-7 (+1e-12
relative); the outputs of the device-resident inner loop and of the
reference-shaped host loop are identical byte for byte.
"""
import csv
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NY, DX = 48, 40, 5.0
INFLOW = [5, 6, 53, 54]
OUTLET = [y * NX + NX - 1 for y in range(NY)]
WET = [y * NX + x for y in range(NY) for x in range(16)]

CONFIG = """! configuration in the reference's mf format
mesh == cartesian
{
  cell count == %(nx)d, %(ny)d
  cell size == %(dx)r, %(dx)r
}
scheme == runge-kutta
{
  method == heun2
  time unit == s
  end time == 120
  step duration == 30
  initial timestep seconds == 1
}
field == z_bed
{
  set == slope
  {
    origin == 0.0, 0.0, 10.0
    slope == -0.02, 0.01
  }
  add == composite
  {
    model == constant
    {
      value == 0.5
      cells == id list
      {
        ids == 300,301,302,348,349,350
      }
    }
  }
}
field == h
{
  set == 2.0
  {
    within == id list
    {
      ids == %(wet)s
    }
  }
}
source term == manning roughness
{
  default shallow n == 0.03
  default deep n == 0.06
}
time series == rainfall
{
  source == inline
  {
    time unit == s
    data ==
    {
      0 == 2e-5
      60 == 2e-5
      90 == 5e-6
      400 == 0
    }
  }
}
boundary == discharge
{
  time series == rain
  {
    values == rainfall
    cells == global
  }
  constant == inflow
  {
    value == 1e-4
    cells == id list
    {
      ids == %(inflow)s
    }
  }
}
boundary == head
{
  constant == outlet
  {
    value == 0.0
    cells == id list
    {
      ids == %(outlet)s
    }
  }
}
output == depth
{
  every == 60
  field == h
  field == stage
}
output == fluxes
{
  every == 120
  at == faces
  field == flux-h
}
measures ==
{
  h-point == gauge
  {
    location == 12.5, 100.0
  }
}
"""


def _say(*a):
    print("[gpu-config-run]", *a, flush=True)


def _inputs():
    x = 0.0 + (np.arange(NX) + 0.5)[None, :] * DX + (np.arange(NY) + 0.5)[:, None] * 0.0
    y = 0.0 + (np.arange(NX) + 0.5)[None, :] * 0.0 + (np.arange(NY) + 0.5)[:, None] * DX
    zb = 10.0 + -0.02 * (x - 0.0) + 0.01 * (y - 0.0)
    zb = zb.ravel()
    zb[[300, 301, 302, 348, 349, 350]] = zb[[300, 301, 302, 348, 349, 350]] * 1.0 + 0.5
    h = np.zeros(NX * NY)
    h[WET] = 2.0
    return zb.reshape(NY, NX), h.reshape(NY, NX)


def _oracle():
    from oracle import oracle as O
    zb, h = _inputs()
    sim = O.Sim(NX, NY, DX, DX, "heun2", 0, 1.0)
    sim.set_bed(zb)
    sim.set_state(h, np.zeros_like(h), np.zeros_like(h))
    sim.set_manning(0.03, 0.06, 0.1, 0.3)
    sim.add_discharge([0.0, 60.0, 90.0, 400.0], [2e-5, 2e-5, 5e-6, 0.0])
    sel = np.zeros(NX * NY, np.uint8)
    sel[INFLOW] = 1
    sim.add_discharge([0.0, 1e12], [1e-4, 1e-4], sel.reshape(NY, NX))
    sel = np.zeros(NX * NY, np.uint8)
    sel[OUTLET] = 1
    sim.add_head([0.0, 1e12], [0.0, 0.0], sel.reshape(NY, NX))
    sim.set_time(0.0, 120.0, 30.0, 1.0)
    return sim


def _run(tmp_path, *opts):
    from mfcm_amd import drive_io
    if not os.path.exists(drive_io.DRIVER):
        from mfcm_amd import build as B
        B.build_driver()
    tmp_path.mkdir(parents=True, exist_ok=True)
    cfg = tmp_path / "run.mf"
    cfg.write_text(CONFIG % {"nx": NX, "ny": NY, "dx": DX, "wet": ",".join(map(str, WET)),
                             "inflow": ",".join(map(str, INFLOW)),
                             "outlet": ",".join(map(str, OUTLET))})
    dump = tmp_path / "dump"
    dump.mkdir()
    cmd = [drive_io.DRIVER, "--quiet", "--dump-state", str(dump), *opts, str(cfg)]
    _say(" ".join(cmd[1:]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    st = dict(l.split(" = ") for l in (dump / "stats.txt").read_text().strip().splitlines())
    h, u, v = (np.fromfile(dump / f"{q}.f64").reshape(NY, NX) for q in "huv")
    return h, u, v, st


def _csv(path):
    with open(path) as f:
        rows = list(csv.reader(f, skipinitialspace=True))
    head = rows[0]
    return head, np.array([[float(x) for x in r if x != ""] for r in rows[1:]])


def test_config_run_matches_oracle(gpu_lib, tmp_path):
    sim = _oracle()
    # oracle step by step, for the gauge (recorded at each step start)
    sim.dt = 1.0
    gauge = []
    for i in range(4):
        gauge.append(sim.get_state(0)[0].ravel()[20 * NX + 2])
        sim.run_steps(i, 1)
    ho, uo, vo = sim.get_state(0)
    acc, rej = sim.stats()

    h, u, v, st = _run(tmp_path / "dev")
    assert (int(st["timesteps"]), int(st["repeated_timesteps"])) == (acc, rej)
    assert np.abs(h - ho).max() <= 1e-9, np.abs(h - ho).max()
    assert np.abs(h * u - ho * uo).max() <= 1e-9
    assert np.abs(h * v - ho * vo).max() <= 1e-9
    np.testing.assert_array_equal(h > 1e-6, ho > 1e-6)
    _say(f"state: {int(np.sum(h != ho))} of {h.size} depths differ in any bit; {acc} + {rej} timesteps")

    out = tmp_path / "dev" / "output"
    names = sorted(os.listdir(out))
    assert names == ["depth_0.000000.csv", "depth_120.000000.csv", "depth_60.000000.csv",
                     "fluxes_0.000000.csv", "fluxes_120.000000.csv", "measures"], names
    head, t = _csv(out / "depth_120.000000.csv")
    assert head[:4] == ["X", "Y", "h", "stage"]
    assert np.abs(t[:, 2] - ho.ravel()).max() <= 5e-7 + 1e-12
    # the stage field is the one of the last step start (t = 90 s)
    head, f = _csv(out / "fluxes_120.000000.csv")
    assert head[:3] == ["X", "Y", "flux-h"] and f.shape[0] == (NX + 1) * NY + NX * (NY + 1)
    hflux = sim.array("hflux")
    assert np.abs(f[:, 2] - hflux).max() <= 5e-7 + 1e-12 * np.abs(hflux).max()
    # before the first timestep the flux field is zero
    _, f0 = _csv(out / "fluxes_0.000000.csv")
    assert np.all(f0[:, 2] == 0.0)

    head, c = _csv(tmp_path / "dev" / "check" / "constants.csv")
    assert head[:5] == ["X", "Y", "z_bed", "dzdx_bed", "dzdy_bed"]
    assert np.abs(c[:, 2] - sim.array("zb")).max() <= 5e-7
    assert np.abs(c[:, 3] - sim.array("dzbdx")).max() <= 5e-7
    assert np.abs(c[:, 4] - sim.array("dzbdy")).max() <= 5e-7

    head, g = _csv(out / "measures" / "gauge.csv")
    assert head[:2] == ["ID", "gauge"]
    np.testing.assert_array_equal(g[:4, 0], [0.0, 30.0, 60.0, 90.0])
    assert np.abs(g[:4, 1] - np.array(gauge)).max() <= 5e-7
    assert np.isnan(g[4, 1])


def test_config_run_host_loop_equals_device_loop(gpu_lib, tmp_path):
    a = _run(tmp_path / "dev")
    b = _run(tmp_path / "host", "--host-loop")
    c = _run(tmp_path / "eager", "--eager")
    for x, y, z in zip(a[:3], b[:3], c[:3]):
        np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(x, z)
    # the device loops also report the terms of their FP64 mass balance
    assert set(b[3]) < set(a[3]) and set(c[3]) == set(a[3])
    assert {k: a[3][k] for k in b[3]} == b[3] and a[3] == c[3]
    vol, vol0, src = (float(a[3][k]) for k in ("volume", "volume0", "source_volume"))
    assert abs(vol - vol0 - src) <= 1e-12 * vol
    for name in ("depth_60.000000.csv", "depth_120.000000.csv", "fluxes_120.000000.csv"):
        assert ((tmp_path / "dev" / "output" / name).read_bytes() ==
                (tmp_path / "host" / "output" / name).read_bytes())


def test_config_run_f32(gpu_lib, tmp_path):
    h, u, v, st = _run(tmp_path / "f32", "--precision", "f32")
    ho, _, _, sto = _run(tmp_path / "f64")
    assert st["precision"] == "f32"
    assert np.all(np.isfinite(h)) and np.all(np.isfinite(u)) and np.all(np.isfinite(v))
    assert abs(float(st["volume"]) - float(sto["volume"])) <= 1e-5 * float(sto["volume"])
    assert np.abs(h - ho).max() <= 0.05
