"""HIP path vs the CPU oracle, through the C ABI (libmfsv_hip.so).

FP64 plain method: the HIP kernels repeat the reference's operation order,
so states agree to the last bits apart from libm-vs-device pow/exp.  The
bar written into each test:
  * one dU/dt evaluation: |k_gpu - k_oracle| <= 1e-12 * (1 + |k|), branch
    codes of every face identical;
  * a run of many timesteps: max |dh|, |d(hu)|, |d(hv)| <= 1e-9 (m, m^2/s),
    wet masks (h > 1e-6) identical, volume relative 1e-10, and the same
    accepted / rejected timestep counts.
"""
import numpy as np
import pytest

from mfcm_amd import scenarios as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-parity]", *a, flush=True)


def _random_state(nx, ny, seed, wet_frac=0.6):
    rng = np.random.default_rng(seed)
    zb = rng.uniform(0.0, 2.0, (ny, nx))
    h = np.where(rng.uniform(size=(ny, nx)) < wet_frac, rng.uniform(0.0, 1.5, (ny, nx)), 0.0)
    u = rng.normal(0, 0.7, (ny, nx)) * (h > 0)
    v = rng.normal(0, 0.7, (ny, nx)) * (h > 0)
    return zb, h, u, v


def _scenario_from_state(nx, ny, dx, dy, zb, h, u, v, flags=0, rain=0.0, head_cols=0, n=0.03):
    z = np.zeros((ny, nx))
    sc = S.Scenario("rand", nx, ny, dx, dy, zb, h, u, v, n, flags=flags, head_cols=head_cols,
                    t_end=120.0, step_duration=30.0)
    if rain:
        sc.rain_t = np.array([0.0, 60.0, 90.0, 400.0])
        sc.rain_v = np.array([rain, rain, 0.3 * rain, 0.0])
    return sc


def _oracle(sc):
    from oracle import oracle as O
    return O.sim_from_scenario(sc)


@pytest.mark.parametrize("nx,ny,flags,seed", [
    (64, 48, 0, 1),
    (70, 40, 0, 11),           # power-of-two spacing: division by multiplication
    (50, 30, S.E7_Y_ORIENT | S.E8_DRY_VELOCITY | S.E9_VELOCITY_CAP, 12),
    (61, 17, 0, 2),            # ragged: one column past a strip, rows not a tile multiple
    (130, 33, S.E7_Y_ORIENT, 3),
    (7, 5, S.E2_WEST_OUTFLOW, 4),
    (1, 9, 0, 5),              # single column
    (13, 1, S.E1_RAIN_GUARD, 6),  # single row
])
def test_dudt_matches_oracle(gpu_lib, nx, ny, flags, seed):
    from mfcm_amd.mfsv import Catchment
    _say(f"dudt nx={nx} ny={ny} flags={flags} seed={seed}")
    zb, h, u, v = _random_state(nx, ny, seed)
    dx, dy = (2.0, 4.0) if seed == 11 else (5.0, 3.0)
    sc = _scenario_from_state(nx, ny, dx, dy, zb, h, u, v, flags=flags, rain=2e-5,
                              head_cols=1 if seed % 2 else 0)
    sim = _oracle(sc)
    sim.start_new_step(30.0)
    dt, t_sub = 0.37, 4.5
    sim.update_dudt(0, t_sub, dt)
    cat = Catchment(sc)
    cat.start_new_step(30.0)
    kh, ku, kv = cat.dudt(t_sub, dt, branches=True)
    for name, g in (("dhdt", kh), ("dudt", ku), ("dvdt", kv)):
        o = sim.array(name).reshape(ny, nx)
        _say(f"  {name}: {int(np.sum(g != o))} of {g.size} values differ in any bit")
        err = np.abs(g - o) / (1.0 + np.abs(o))
        assert np.all(np.isfinite(g)), name
        assert err.max() <= 1e-12, (name, err.max(), np.unravel_index(err.argmax(), err.shape))
    br = sim.array("branch")
    nxf = (nx + 1) * ny
    np.testing.assert_array_equal(cat.download("branch_x").ravel(), br[:nxf])
    np.testing.assert_array_equal(cat.download("branch_y").ravel(), br[nxf:])
    cat.close()


def _compare_states(cat, sim, tol=1e-9):
    hg, ug, vg = cat.state()
    ho, uo, vo = sim.get_state(0)
    assert np.abs(hg - ho).max() <= tol, np.abs(hg - ho).max()
    assert np.abs(hg * ug - ho * uo).max() <= tol, np.abs(hg * ug - ho * uo).max()
    assert np.abs(hg * vg - ho * vo).max() <= tol, np.abs(hg * vg - ho * vo).max()
    np.testing.assert_array_equal(hg > 1e-6, ho > 1e-6)
    return np.abs(hg - ho).max()


@pytest.mark.parametrize("case", ["dambreak_parity", "dambreak_e7", "hillslope_rain_prewet_parity",
                                  "hillslope_rain_physics", "storm_catchment_physics"])
def test_run_matches_oracle(gpu_lib, case):
    from mfcm_amd.mfsv import Catchment
    if case.startswith("dambreak"):
        sc = S.dam_break_bump(64, 48)
        sc.t_end = 60.0 if case.endswith("parity") else 120.0
        if case.endswith("e7"):
            sc.flags |= S.E7_Y_ORIENT
    elif case == "storm_catchment_physics":
        sc = S.storm_catchment(seed=1, n=64, dx=2.0, t_end=180.0)
    elif case == "hillslope_rain_prewet_parity":
        sc = S.hillslope_rain(seed=0, n=64, dx=5.0, prewet=0.002, t_end=300.0)
    else:
        sc = S.hillslope_rain(seed=0, n=64, dx=5.0, physics=True, t_end=240.0)
        sc.flags |= S.E7_Y_ORIENT
    _say(f"run {case} {sc.nx}x{sc.ny} t_end={sc.t_end} flags={sc.flags}")
    sim = _oracle(sc)
    sim.solve()
    cat = Catchment(sc)
    cat.solve()
    err = _compare_states(cat, sim)
    acc, rej = sim.stats()
    assert (cat.n_acc, cat.n_rej) == (acc, rej)
    vg, vo = cat.volume(), sim.volume(0)
    assert abs(vg - vo) <= 1e-10 * max(abs(vo), 1e-30)
    _say(f"  ok: {acc} accepted, {rej} rejected, max|dh|={err:.3g}")
    cat.close()


@pytest.mark.parametrize("graph", [False, True])
def test_device_controller_equals_host_loop(gpu_lib, graph):
    """The device controller (mfsv_run) repeats the host loop bit for bit."""
    from mfcm_amd.mfsv import Catchment
    sc = S.hillslope_rain(seed=4, n=96, dx=5.0, prewet=0.002, t_end=240.0)
    _say(f"device controller graph={graph}")
    a = Catchment(sc)
    a.solve()
    b = Catchment(sc)
    b.run_begin()
    st = b.run(graph=graph)
    assert st["done"] == 1
    assert (st["accepted"], st["rejected"]) == (a.n_acc, a.n_rej)
    for x, y in zip(a.state(), b.state()):
        np.testing.assert_array_equal(x, y)
    a.close()
    b.close()


def test_device_run_matches_oracle_and_mass(gpu_lib):
    from mfcm_amd.mfsv import Catchment
    sc = S.hillslope_rain(seed=5, n=80, dx=5.0, physics=True, t_end=180.0)
    sc.flags |= S.E7_Y_ORIENT
    _say("device run vs oracle, physics flags")
    sim = _oracle(sc)
    sim.solve()
    cat = Catchment(sc)
    cat.run_begin()
    st = cat.run()
    _compare_states(cat, sim)
    # mass balance: volume change == integrated source/boundary rate (FP64)
    resid = st["volume"] - st["volume0"] - st["source_volume"]
    assert abs(resid) <= 1e-9 * max(st["volume"], 1.0), resid
    cat.close()


def test_hillslope_256_matches_oracle(gpu_lib):
    """The CPU-runnable configuration at full size (256^2, dx 5 m, physics
    flags, CFL 0.5) over its first 300 s: FP64 GPU vs the oracle."""
    from mfcm_amd.mfsv import Catchment
    sc = S.hillslope_rain(seed=0, n=256, dx=5.0, physics=True, t_end=300.0)
    _say(f"hillslope 256^2 300 s flags={sc.flags}")
    sim = _oracle(sc)
    sim.solve()
    cat = Catchment(sc)
    cat.run_begin()
    st = cat.run()
    _compare_states(cat, sim)
    acc, rej = sim.stats()
    assert (st["accepted"], st["rejected"]) == (acc, rej)
    vg, vo = cat.volume(), sim.volume(0)
    assert abs(vg - vo) <= 1e-10 * max(abs(vo), 1e-30)
    cat.close()


def _with_sources(sc, seed, kinds):
    """Switch on the further source terms / stage boundary on a scenario."""
    rng = np.random.default_rng(seed)
    ny, nx = sc.ny, sc.nx
    if "eloss" in kinds:
        sc.energy_loss = (rng.uniform(0.0, 0.05, (ny, nx)), rng.uniform(0.0, 0.05, (ny, nx)))
    if "infil" in kinds:
        sc.infiltration = (rng.uniform(0.0, 2e-3, (ny, nx)), rng.uniform(0.0, 5e-3, (ny, nx)))
    if "eddy" in kinds:
        sc.eddy = {"mu": rng.uniform(0.0, 0.3, (ny, nx))}
    if "smag" in kinds:
        sc.eddy = {"smagorinsky": (0.05, 0.5)}
    if "stage" in kinds:
        sc.stage_t = np.array([0.0, 1e6])
        lvl = float(np.median(sc.zb)) + 0.4
        sc.stage_v = np.array([lvl, lvl])
        sel = np.zeros((ny, nx), bool)
        sel[:, -1] = True
        sc.stage_sel = sel
    return sc


@pytest.mark.parametrize("nx,ny,dx,flags,kinds,seed", [
    (64, 48, 5.0, 0, ("eloss", "infil", "eddy", "stage"), 21),
    (70, 40, 2.0, S.E7_Y_ORIENT, ("eloss", "infil", "smag", "stage"), 22),   # power-of-two spacing
    (61, 17, 3.0, S.E7_Y_ORIENT | S.E8_DRY_VELOCITY, ("smag",), 23),
    (13, 9, 4.0, 0, ("eddy", "infil"), 24),
])
def test_dudt_further_sources_match_oracle(gpu_lib, nx, ny, dx, flags, kinds, seed):
    """Energy loss, infiltration (and the capacity it consumes), eddy
    viscosity (per-cell and Smagorinsky) and the stage boundary, one dU/dt
    evaluation: HIP vs oracle."""
    from mfcm_amd.mfsv import Catchment
    _say(f"dudt further sources nx={nx} ny={ny} dx={dx} flags={flags} {kinds}")
    zb, h, u, v = _random_state(nx, ny, seed)
    sc = _scenario_from_state(nx, ny, dx, dx, zb, h, u, v, flags=flags, rain=2e-5)
    sc = _with_sources(sc, seed, kinds)
    sim = _oracle(sc)
    sim.start_new_step(30.0)
    dt, t_sub = 0.37, 4.5
    sim.update_dudt(0, t_sub, dt)
    cat = Catchment(sc)
    cat.start_new_step(30.0)
    kh, ku, kv = cat.dudt(t_sub, dt)
    for name, g in (("dhdt", kh), ("dudt", ku), ("dvdt", kv)):
        o = sim.array(name).reshape(ny, nx)
        err = np.abs(g - o) / (1.0 + np.abs(o))
        _say(f"  {name}: {int(np.sum(g != o))} of {g.size} values differ in any bit")
        assert np.all(np.isfinite(g)), name
        assert err.max() <= 1e-12, (name, err.max(), np.unravel_index(err.argmax(), err.shape))
    if "infil" in kinds:
        np.testing.assert_array_equal(cat.download("infil_cap").ravel(), sim.array("infil_capacity"))
    cat.close()


@pytest.mark.parametrize("graph", [False, True])
def test_run_further_sources_matches_oracle(gpu_lib, graph):
    """A dam break with infiltration, energy loss, Smagorinsky eddy viscosity
    and a constant stage boundary on the east column, run to 60 s: the
    device-controlled run vs the oracle (state, counts, remaining capacity)."""
    from mfcm_amd.mfsv import Catchment
    sc = S.dam_break_bump(64, 48)
    sc.t_end = 60.0
    sc = _with_sources(sc, 31, ("eloss", "infil", "smag", "stage"))
    _say(f"run further sources graph={graph}")
    sim = _oracle(sc)
    sim.solve()
    cat = Catchment(sc)
    cat.run_begin()
    st = cat.run(graph=graph)
    _compare_states(cat, sim)
    acc, rej = sim.stats()
    assert (st["accepted"], st["rejected"]) == (acc, rej)
    cap_g = cat.download("infil_cap").ravel()
    assert np.abs(cap_g - sim.array("infil_capacity")).max() <= 1e-12
    resid = st["volume"] - st["volume0"] - st["source_volume"]
    assert abs(resid) <= 1e-9 * max(st["volume"], 1.0), resid
    cat.close()


@pytest.mark.parametrize("prec_name", ["f64", "f32"])
@pytest.mark.parametrize("storm", [False, True])
def test_per_cell_boundaries_equal_uniform(gpu_lib, prec_name, storm):
    """Per-cell discharge and head endpoints (what the configuration path's
    selections upload) give the same states, bit for bit, as the uniform
    rain series and the head band they restate: the per-cell values are the
    same numbers rounded the same way, so any difference is a kernel reading
    the wrong word (the FP32 form loads these fields ahead of the row
    prefetch, from a stand-in address when a field is off)."""
    import ctypes as C
    from mfcm_amd import mfsv as M
    sc = S.hillslope_rain(seed=6, n=70, dx=5.0, prewet=0.002, t_end=180.0)
    sc.ny = 45
    for a in ("zb", "h0", "u0", "v0"):
        setattr(sc, a, getattr(sc, a)[:45].copy())
    sc.head_cols = 3
    if storm:
        sc.storm = S.drifting_storm(9, sc.nx, sc.ny, sc.dx)
        sc.flags |= S.E4_STORM
    prec = M.F32 if prec_name == "f32" else M.F64
    _say(f"per-cell boundaries {prec_name} storm={storm} {sc.nx}x{sc.ny}")
    a = M.Catchment(sc, precision=prec)
    a.solve()
    b = M.Catchment(sc, precision=prec)
    b.params.discharge, b.params.head = 2, 2
    band = np.zeros((sc.ny, sc.nx), bool)
    band[:, :sc.head_cols] = True
    sd = float(sc.step_duration)
    b.dt = float(sc.initial_dt)
    for i in range(b.nsteps):
        t = 0.0 + i * sd
        q0, q1, hb0, hb1 = b.forcing_at(t)
        b.upload("qbdy0", np.full((sc.ny, sc.nx), q0))
        b.upload("qbdy1", np.full((sc.ny, sc.nx), q1))
        b.upload("hbdy0", np.where(band, hb0, np.nan))
        b.upload("hbdy1", np.where(band, hb1, np.nan))
        if i == 0:
            b._check(M.lib().mfsv_set_params(b._h, C.byref(b.params)))
        b._check(M.lib().mfsv_start_new_step(b._h, t, q0, q1, hb0, hb1))
        b.step()
    assert (a.n_acc, a.n_rej) == (b.n_acc, b.n_rej)
    for x, y in zip(a.state(), b.state()):
        np.testing.assert_array_equal(x, y)
    _say(f"  ok: {a.n_acc} accepted, {a.n_rej} rejected, wet cells {int(np.sum(a.state()[0] > 1e-6))}")
    a.close()
    b.close()


def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("nx,ny,flags,seed", [
    (64, 48, 0, 21),
    (61, 17, S.E7_Y_ORIENT | S.E8_DRY_VELOCITY | S.E9_VELOCITY_CAP, 22),
    (130, 33, S.E7_Y_ORIENT | S.E1_RAIN_GUARD, 23),
])
def test_dudt_f32_matches_oracle(gpu_lib, nx, ny, flags, seed):
    """The FP32 precision form's dU/dt against the FP64 oracle on the same
    (FP32-representable) inputs, power-of-two spacing: FP32 rounding only,
    |k32 - k64| <= 1e-5 (1 + |k64|) everywhere and every value finite."""
    from mfcm_amd.mfsv import Catchment, F32
    zb, h, u, v = (_f32(a) for a in _random_state(nx, ny, seed))
    sc = _scenario_from_state(nx, ny, 2.0, 4.0, zb, h, u, v, flags=flags, rain=2e-5,
                              head_cols=1 if seed % 2 else 0)
    _say(f"dudt f32 nx={nx} ny={ny} flags={flags} seed={seed}")
    sim = _oracle(sc)
    sim.start_new_step(30.0)
    dt, t_sub = 0.375, 4.5
    sim.update_dudt(0, t_sub, dt)
    cat = Catchment(sc, precision=F32)
    cat.start_new_step(30.0)
    kh, ku, kv = cat.dudt(t_sub, dt)
    for name, g in (("dhdt", kh), ("dudt", ku), ("dvdt", kv)):
        o = sim.array(name).reshape(ny, nx)
        assert np.all(np.isfinite(g)), name
        err = np.abs(g - o) / (1.0 + np.abs(o))
        _say(f"  {name}: max rel err {err.max():.3g}")
        assert err.max() <= 1e-5, (name, err.max(), np.unravel_index(err.argmax(), err.shape))
    cat.close()


@pytest.mark.parametrize("e7", [False, True])
def test_run_f32_close_to_oracle(gpu_lib, e7):
    """A dam break around a bump (flow along both axes) for 30 s in the FP32
    form against the FP64 oracle: the depth stays within 5e-4 m everywhere
    (FP32 rounding over ~100 timesteps; an error in the y treatment moves
    the front by whole cells), and the volume within 1e-6."""
    from mfcm_amd.mfsv import Catchment, F32
    sc = S.dam_break_bump(64, 48, dx=4.0, dy=4.0)
    sc.zb = _f32(sc.zb)
    sc.t_end = 30.0
    if e7:
        sc.flags |= S.E7_Y_ORIENT
    _say(f"run f32 dam break e7={e7}")
    sim = _oracle(sc)
    sim.solve()
    cat = Catchment(sc, precision=F32)
    cat.solve()
    hg = cat.state()[0]
    ho = sim.get_state(0)[0]
    err = np.abs(hg - ho).max()
    _say(f"  max |dh| {err:.3g} m; timesteps {cat.n_acc}+{cat.n_rej} vs {sim.stats()}")
    assert err <= 5e-4, err
    vg, vo = cat.volume(), sim.volume(0)
    assert abs(vg - vo) <= 1e-6 * vo, (vg, vo)
    cat.close()
