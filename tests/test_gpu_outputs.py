"""Output-only fields of the solver (what MultiFieldOutputFile and the check
files write) and per-cell Manning parameters: HIP vs the CPU oracle through
the C ABI.

Bar, as for the dU/dt parity tests: every value within 1e-12 * (1 + |ref|)
for one evaluation and 1e-9 * (1 + |ref|) after a run (FP64 plain method; in
practice bit-identical), face order and count as the reference's
Cartesian2DMesh (x-faces, then y-faces).
"""
import numpy as np
import pytest

from mfcm_amd import scenarios as S

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print("[gpu-outputs]", *a, flush=True)


def _close(name, g, o, tol=TOL):
    g = np.asarray(g).ravel()
    o = np.asarray(o).ravel()
    assert g.shape == o.shape, (name, g.shape, o.shape)
    both_nan = np.isnan(g) & np.isnan(o)
    err = np.where(both_nan, 0.0, np.abs(g - o) / (1.0 + np.abs(o)))
    _say(f"  {name}: {int(np.sum((g != o) & ~both_nan))} of {g.size} values differ in any bit")
    assert not np.isnan(err).any(), (name, "NaN in one result only")
    assert err.max() <= tol, (name, err.max(), int(err.argmax()))


def _random_scenario(nx, ny, dx, dy, seed, flags, manning_fields):
    rng = np.random.default_rng(seed)
    zb = rng.uniform(0.0, 2.0, (ny, nx))
    h = np.where(rng.uniform(size=(ny, nx)) < 0.6, rng.uniform(0.0, 1.5, (ny, nx)), 0.0)
    u = rng.normal(0, 0.7, (ny, nx)) * (h > 0)
    v = rng.normal(0, 0.7, (ny, nx)) * (h > 0)
    sc = S.Scenario("rand", nx, ny, dx, dy, zb, h, u, v, 0.03, flags=flags, t_end=120.0,
                    step_duration=30.0)
    sc.rain_t = np.array([0.0, 60.0, 90.0, 400.0])
    sc.rain_v = np.array([2e-5, 2e-5, 6e-6, 0.0])
    if manning_fields:
        sc.manning_fields = (rng.uniform(0.02, 0.05, (ny, nx)), rng.uniform(0.02, 0.2, (ny, nx)),
                             rng.uniform(0.05, 0.2, (ny, nx)), rng.uniform(0.3, 0.6, (ny, nx)))
    return sc


@pytest.mark.parametrize("nx,ny,dx,dy,flags,mf,seed", [
    (64, 48, 5.0, 3.0, 0, True, 1),
    (70, 40, 2.0, 4.0, S.E7_Y_ORIENT, True, 2),      # power-of-two spacing (difference form)
    (61, 17, 3.0, 3.0, S.E2_WEST_OUTFLOW | S.E8_DRY_VELOCITY, False, 3),
    (7, 5, 1.0, 1.0, 0, True, 4),
    (1, 9, 5.0, 5.0, 0, False, 5),
])
def test_outputs_of_current_state_match_oracle(gpu_lib, nx, ny, dx, dy, flags, mf, seed):
    """Face fluxes, Manning n, friction slope and bed slopes of U0, and the
    dU/dt with per-cell n_shallow/n_deep/d_shallow/d_deep."""
    from oracle import oracle as O
    from mfcm_amd.mfsv import Catchment
    _say(f"outputs of U0 nx={nx} ny={ny} dx={dx} dy={dy} flags={flags} manning_fields={mf}")
    sc = _random_scenario(nx, ny, dx, dy, seed, flags, mf)
    sim = O.sim_from_scenario(sc)
    sim.start_new_step(30.0)
    dt, t_sub = 0.37, 4.5
    sim.update_dudt(0, t_sub, dt)
    cat = Catchment(sc)
    cat.start_new_step(30.0)
    kh, ku, kv = cat.dudt(t_sub, dt)
    for name, g in (("dhdt", kh), ("dudt", ku), ("dvdt", kv)):
        _close(name, g, sim.array(name))
    d = cat.diagnose(0, t_sub, dt)
    for g, o in (("flux_h", "hflux"), ("flux_u", "uflux"), ("flux_v", "vflux"), ("flux_z", "zflux"),
                 ("mannings_n", "mannings_n"), ("friction_slope", "friction_slope"),
                 ("dzdx_bed", "dzbdx"), ("dzdy_bed", "dzbdy")):
        _close(g, d[g], sim.array(o))
    # the evaluation has no side effects on the derivative fields' source
    kh2, ku2, kv2 = cat.dudt(t_sub, dt)
    np.testing.assert_array_equal(kh2, kh)
    cat.close()


def test_outputs_do_not_consume_infiltration(gpu_lib):
    from mfcm_amd.mfsv import Catchment
    sc = _random_scenario(40, 30, 5.0, 5.0, 7, 0, False)
    rng = np.random.default_rng(7)
    sc.infiltration = (rng.uniform(0.0, 2e-3, (30, 40)), rng.uniform(0.0, 5e-3, (30, 40)))
    cat = Catchment(sc)
    cat.start_new_step(0.0)
    cap0 = cat.download("infil_cap")
    cat.diagnose(0, 0.0, 0.5)
    np.testing.assert_array_equal(cat.download("infil_cap"), cap0)
    cat.close()


@pytest.mark.parametrize("scheme,graph", [("heun2", True), ("heun2", False), ("euler", True),
                                          ("classic", True)])
def test_outputs_of_last_stage_match_oracle(gpu_lib, scheme, graph):
    """After a device-controlled run: the fluxes, Manning n and friction slope
    of the last RK stage's input -- what the reference's fields hold when it
    writes its outputs -- and the stage field z + h of the final state."""
    from oracle import oracle as O
    from mfcm_amd.mfsv import Catchment
    _say(f"outputs of the last stage input scheme={scheme} graph={graph}")
    sc = S.dam_break_bump(48, 40)
    sc.t_end = 30.0
    sc.scheme = scheme
    rng = np.random.default_rng(9)
    sc.manning_fields = (np.full((40, 48), 0.03), rng.uniform(0.03, 0.1, (40, 48)),
                         np.full((40, 48), 0.1), np.full((40, 48), 0.4))
    sim = O.sim_from_scenario(sc)
    sim.solve()
    cat = Catchment(sc)
    cat.run_begin()
    st = cat.run(graph=graph)
    assert (st["accepted"], st["rejected"]) == sim.stats()
    d = cat.diagnose(1)
    for g, o in (("flux_h", "hflux"), ("flux_u", "uflux"), ("flux_v", "vflux"), ("flux_z", "zflux"),
                 ("mannings_n", "mannings_n"), ("friction_slope", "friction_slope")):
        # states after a run agree to 1e-9 (the run bar), their fluxes likewise
        _close(g, d[g], sim.array(o), tol=1e-9)
    h, _, _ = cat.state()
    stage = cat.stage_field()
    np.testing.assert_array_equal(stage, sc.zb + h)
    # the run continues unchanged after the evaluation
    st2 = cat.stats()
    assert st2["t_local"] == st["t_local"] and st2["dt"] == st["dt"]
    cat.close()


def test_last_stage_outputs_need_a_timestep(gpu_lib):
    from mfcm_amd.mfsv import Catchment, MfsvError
    cat = Catchment(S.dam_break_bump(16, 16))
    with pytest.raises(MfsvError):
        cat.diagnose(1)
    cat.close()


@pytest.mark.parametrize("nx,ny,dx,dy,seed", [(64, 48, 5.0, 3.0, 11), (70, 33, 2.0, 4.0, 12),
                                              (1, 9, 5.0, 5.0, 13), (13, 1, 3.0, 3.0, 14)])
def test_spatial_derivative_matches_oracle(gpu_lib, nx, ny, dx, dy, seed):
    """SpatialDerivativeOperator with Minmod3 on a field with coded-out (NaN)
    cells, along X and along the reference's Y: identical to the oracle's
    slope passes, NaN where the oracle has NaN."""
    from oracle import oracle as O
    from mfcm_amd.mfsv import Catchment
    _say(f"spatial derivative nx={nx} ny={ny} dx={dx} dy={dy}")
    rng = np.random.default_rng(seed)
    f = rng.normal(0.0, 1.0, (ny, nx))
    f[rng.uniform(size=(ny, nx)) < 0.1] = np.nan
    f[rng.uniform(size=(ny, nx)) < 0.1] = 0.25        # ties
    sc = S.Scenario("sd", nx, ny, dx, dy, np.zeros((ny, nx)), np.zeros((ny, nx)),
                    np.zeros((ny, nx)), np.zeros((ny, nx)), 0.03)
    cat = Catchment(sc)
    gx = cat.spatial_derivative(f, 0)
    gy = cat.spatial_derivative(f, 1)
    np.testing.assert_array_equal(gx, O.slope_x(f, dx))
    np.testing.assert_array_equal(gy, O.slope_y(f, dy, 0))
    cat.close()


@pytest.mark.parametrize("scheme", ["heun2", "classic"])
def test_control_number_of_each_state(gpu_lib, scheme):
    """control_number(state, dt) of U0, the intermediate RK rows and the
    result of one timestep, against the oracle's after the same timestep;
    the result's equals the value the fused timestep returned."""
    from oracle import oracle as O
    from mfcm_amd.mfsv import Catchment
    sc = S.dam_break_bump(48, 40)
    sc.scheme = scheme
    sim = O.sim_from_scenario(sc)
    sim.start_new_step(0.0)
    dt = 0.6
    sim.do_timestep(0.0, dt)
    cat = Catchment(sc)
    cat.start_new_step(0.0)
    cn_fused = cat.control_timestep(0.0, dt)
    n = cat.nstages
    for s in range(n + 1):
        g = cat.control_number(s, dt)
        o = sim.control_number(s, dt)
        _say(f"  {scheme} state {s}: {g!r} vs {o!r}")
        assert g == pytest.approx(o, rel=1e-14), (s, g, o)
    assert cn_fused == pytest.approx(cat.control_number(n, dt), rel=1e-15)
    cat.close()
