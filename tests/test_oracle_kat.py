"""Known-answer and invariant tests that pin the CPU oracle.

The reference ships no tests, golden vectors or fixtures for this path, so
every oracle function is pinned here by what the mathematics fixes:
hand-computed face fluxes for each branch, closed-form Minmod3 cases, exact
invariants (a lake at rest stays at rest, wall faces cancel, linear data has
its exact slope), closed forms for the sources, the Butcher tables' order
conditions, the controller's table and the landing rules.  A dropped term, a
wrong sign, a swapped neighbour or a transposed operand in the oracle fails
at least one of these.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

G = 9.81


def cell(**kw):
    d = dict(h=0.0, u=0.0, v=0.0, zb=0.0, dzbdx=0.0, dzbdy=0.0, dhdx=0.0, dudx=0.0,
             dvdx=0.0, dhdy=0.0, dudy=0.0, dvdy=0.0)
    d.update(kw)
    return d


# ---------------------------------------------------------------- Minmod3
@pytest.mark.parametrize("ul,uc,ur,d,want", [
    (0.0, 1.0, 2.0, 1.0, 1.0),        # linear: central difference
    (0.0, 1.0, 1.1, 1.0, 0.2),        # limited by the steep side: 2*(1.1-1)
    (0.0, 1.0, 0.0, 1.0, 0.0),        # extremum
    (1.0, 1.0, 3.0, 2.0, 0.0),        # flat on one side: sign(0) = 0
    (3.0, 2.0, 0.0, 0.5, -3.0),       # decreasing: outer -min(4, 8), central (0-3)/1 = -3
    (0.0, 0.25, 2.0, 1.0, 0.5),       # 2*0.25 < central 1.0
])
def test_minmod3_interior(ul, uc, ur, d, want):
    assert O.minmod3(ul, d, uc, d, ur) == pytest.approx(want, rel=0, abs=1e-15)


def test_minmod3_edges_are_one_sided():
    # missing neighbour = the cell itself at distance 0 (Cartesian2DMesh.hpp:477-520)
    assert O.minmod3(5.0, 0.0, 5.0, 2.0, 7.0) == 1.0
    assert O.minmod3(7.0, 2.0, 5.0, 0.0, 5.0) == -1.0
    assert O.minmod3(5.0, 0.0, 5.0, 0.0, 5.0) == 0.0


def test_minmod3_nan_neighbour_semantics():
    # lhs NaN -> outer = rhs; central NaN: (b < a ? b : a) keeps |rhs|, sign(NaN) = 0
    assert O.minmod3(float("nan"), 1.0, 1.0, 1.0, 2.0) == 1.0


def test_minmod3_odd_symmetry():
    rng = np.random.default_rng(0)
    for _ in range(500):
        a, b, c = rng.normal(size=3)
        assert O.minmod3(-a, 1.5, -b, 1.5, -c) == -O.minmod3(a, 1.5, b, 1.5, c)


def test_slopes_of_linear_fields_are_exact():
    ny, nx, dx, dy = 7, 9, 2.0, 4.0
    y, x = np.mgrid[0:ny, 0:nx].astype(float)
    f = 3.0 + 0.5 * x * dx - 0.25 * y * dy          # exactly representable steps
    sx = O.slope_x(f, dx)
    assert np.all(sx == 0.5)
    # reference orientation: north (row + 1) is "left" -> minus d/dy
    assert np.all(O.slope_y(f, dy) == 0.25)
    assert np.all(O.slope_y(f, dy, flags=O.E7_Y_ORIENT) == -0.25)


# ---------------------------------------------------------------- face flux
def test_face_still_equal_depths_is_zero_flux():
    L = cell(h=1.0)
    out = O.face(L, L, 0, 0, 1.0)
    assert out == [0.0, 0.0, 0.0, 0.0, 1.0]


def test_face_still_unequal_depths_jump_term():
    # branch 1 with zero velocity: a = 0 (sign(0) = 0), u-flux = g (hL - hR) / 2
    out = O.face(cell(h=2.0), cell(h=1.0), 0, 0, 1.0)
    assert out[0] == 0.0 and out[4] == 1.0
    assert out[1] == pytest.approx(0.5 * G * (2.0 - 1.0), abs=1e-15)
    assert out[3] == 0.0


def test_face_uniform_flow():
    c = math.sqrt(G)
    out = O.face(cell(h=1.0, u=1.0), cell(h=1.0, u=1.0), 0, 0, 1.0)
    assert out[0] == pytest.approx(1.0, abs=1e-15)              # h u
    assert out[1] == pytest.approx(0.5, abs=1e-15)              # u^2/2 ... central part
    assert out[2] == 0.0
    # the same on a y-face transports v
    out = O.face(cell(h=1.0, v=1.0), cell(h=1.0, v=1.0), 0, 1, 1.0)
    assert out[0] == pytest.approx(1.0) and out[2] == pytest.approx(0.5) and out[1] == 0.0
    del c


def test_face_tangential_momentum_flux():
    # v on an x-face is transported as v * u (FluxKernel.cpp:220-223)
    out = O.face(cell(h=1.0, u=2.0, v=3.0), cell(h=1.0, u=2.0, v=3.0), 0, 0, 1.0)
    assert out[2] == pytest.approx(3.0 * 2.0, abs=1e-14)


def test_face_branch_dry_right_higher():
    out = O.face(cell(h=0.1, zb=0.0), cell(h=0.0, zb=1.0), 0, 0, 1.0)
    assert out[4] == 2.25
    assert out[0] == 0.0
    assert out[1] == pytest.approx(0.5 * G * 0.1)
    assert out[3] == pytest.approx(-0.5 * G * 0.1)


def test_face_branch_dry_left_higher():
    out = O.face(cell(h=0.0, zb=1.0), cell(h=0.1, zb=0.0), 0, 0, 1.0)
    assert out[4] == 2.75
    assert out[1] == pytest.approx(-0.5 * G * 0.1)
    assert out[3] == pytest.approx(0.5 * G * 0.1)


def test_face_branch_left_wet_only():
    out = O.face(cell(h=2.0, u=0.5, zb=0.0), cell(h=0.0, zb=1.0), 0, 0, 1.0)
    a = 0.5 + math.sqrt(G * 2.0)
    assert out[4] == 3.0
    assert out[0] == pytest.approx(0.5 * (2.0 * 0.5) + 0.5 * a * 2.0, rel=1e-15)
    assert out[1] == pytest.approx(0.5 * (0.5 * 0.25 + G * 2.0) + 0.5 * a * 0.5, rel=1e-15)
    assert out[3] == 0.0                                           # h_p / dx with h_p = 0


def test_face_branch_right_wet_only_uses_plus_g_h():
    out = O.face(cell(h=0.0, zb=1.0), cell(h=2.0, u=-0.5, zb=0.0), 0, 0, 2.0)
    a = 0.5 + math.sqrt(G * 2.0)
    assert out[4] == 4.0
    assert out[0] == pytest.approx(0.5 * (-1.0) - 0.5 * a * 2.0, rel=1e-15)
    # FluxKernel.cpp:277-278: + g h_p on the plus side in this branch
    assert out[1] == pytest.approx(0.5 * (0.125 + G * 2.0) - 0.5 * a * (-0.5), rel=1e-15)
    assert out[3] == 0.0


def test_face_wall_ghost_and_cancellation():
    R = cell(h=1.0, zb=3.0)
    out = O.face(R, R, -1, 0, 1.0)           # west wall: ghost bed zb + 2h
    assert out[4] == 2.75
    assert out[1] == pytest.approx(-0.5 * G) and out[3] == pytest.approx(0.5 * G)
    assert out[1] + out[3] == 0.0            # wall force cancels in the cell pass
    out = O.face(R, R, 1, 1, 1.0)            # north wall
    assert out[4] == 2.25
    assert out[2] == pytest.approx(0.5 * G) and out[3] == pytest.approx(-0.5 * G)


def test_face_coded_out_cells():
    R = cell(h=1.0, zb=3.0)
    nan = cell(zb=float("nan"))
    # NaN on the left of an interior face behaves as a west wall of R
    assert O.face(nan, R, 0, 0, 1.0) == O.face(R, R, -1, 0, 1.0)
    # NaN on both sides, or NaN against a mesh edge: zero flux, branch 0
    assert O.face(nan, nan, 0, 0, 1.0) == [0.0] * 5
    assert O.face(nan, nan, 1, 0, 1.0) == [0.0] * 5


# ---------------------------------------------------------------- whole-step invariants
def _lake(nx=12, ny=9, h=1.0, dx=3.0, dy=2.0, flags=0):
    sim = O.Sim(nx, ny, dx, dy, "heun2", flags)
    sim.set_bed(np.zeros((ny, nx)))
    sim.set_state(np.full((ny, nx), h))
    sim.set_manning(0.03)
    sim.set_time(0.0, 600.0, 60.0, 1.0)
    return sim


@pytest.mark.parametrize("flags", [0, O.E7_Y_ORIENT])
def test_lake_at_rest_stays_at_rest(flags):
    sim = _lake(flags=flags)
    sim.update_dudt(0, 0.0, 0.5)
    for k in ("dhdt", "dudt", "dvdt"):
        assert np.all(sim.array(k) == 0.0), k


def test_closed_basin_flux_sum_telescopes():
    rng = np.random.default_rng(1)
    nx, ny = 20, 15
    sim = O.Sim(nx, ny, 2.0, 2.0, "heun2", O.E7_Y_ORIENT)
    sim.set_bed(np.zeros((ny, nx)))
    h = 1.0 + 0.1 * rng.random((ny, nx))
    sim.set_state(h, 0.2 * rng.normal(size=(ny, nx)), 0.2 * rng.normal(size=(ny, nx)))
    sim.update_dudt(0, 0.0, 0.1)
    dh = sim.array("dhdt")
    # wall faces of a wet basin on a flat bed carry no water
    hf = sim.array("hflux")
    nxf = (nx + 1) * ny
    west = hf[:nxf].reshape(ny, nx + 1)[:, 0]
    east = hf[:nxf].reshape(ny, nx + 1)[:, nx]
    assert np.all(west == 0.0) and np.all(east == 0.0)
    assert abs(dh.sum()) <= 1e-12 * np.abs(dh).sum()


def test_xy_transpose_with_e7():
    rng = np.random.default_rng(2)
    nx, ny = 11, 7
    zb = rng.random((ny, nx))
    h = 0.5 + rng.random((ny, nx))
    u, v = 0.3 * rng.normal(size=(ny, nx)), 0.3 * rng.normal(size=(ny, nx))
    a = O.Sim(nx, ny, 2.0, 3.0, "heun2", O.E7_Y_ORIENT)
    a.set_bed(zb); a.set_state(h, u, v); a.set_manning(0.03)
    a.update_dudt(0, 0.0, 0.1)
    b = O.Sim(ny, nx, 3.0, 2.0, "heun2", O.E7_Y_ORIENT)
    b.set_bed(zb.T); b.set_state(h.T, v.T, u.T); b.set_manning(0.03)
    b.update_dudt(0, 0.0, 0.1)
    np.testing.assert_array_equal(a.array("dhdt").reshape(ny, nx), b.array("dhdt").reshape(nx, ny).T)
    np.testing.assert_array_equal(a.array("dudt").reshape(ny, nx), b.array("dvdt").reshape(nx, ny).T)


def test_reference_y_orientation_is_not_symmetric():
    """Without E7 the y pressure term has the reference's sign (quirk): a
    y-dam-break differs from the transposed x-dam-break."""
    nx = ny = 10
    h = 1.0 + 0.125 * np.arange(nx)[None, :] + np.zeros((ny, 1))   # linear ramp along x
    a = O.Sim(nx, ny, 1.0, 1.0); a.set_bed(np.zeros((ny, nx))); a.set_state(h)
    a.update_dudt(0, 0.0, 0.1)
    b = O.Sim(ny, nx, 1.0, 1.0); b.set_bed(np.zeros((nx, ny))); b.set_state(h.T)
    b.update_dudt(0, 0.0, 0.1)
    du = a.array("dudt").reshape(ny, nx)
    dv = b.array("dvdt").reshape(nx, ny).T
    # interior: pressure term -g dh/dx along x, +g dh/dy along y (wrong sign)
    assert np.allclose(du[2:-2, 2:-2], -G * 0.125, rtol=1e-14)
    assert np.allclose(dv[2:-2, 2:-2], +G * 0.125, rtol=1e-14)
    c = O.Sim(ny, nx, 1.0, 1.0, "heun2", O.E7_Y_ORIENT)
    c.set_bed(np.zeros((nx, ny))); c.set_state(h.T)
    c.update_dudt(0, 0.0, 0.1)
    np.testing.assert_array_equal(c.array("dvdt").reshape(nx, ny).T, du)


# ---------------------------------------------------------------- sources
def _single(h, u=0.0, v=0.0, n=0.03, dt=0.5):
    sim = O.Sim(1, 1, 10.0, 10.0)
    sim.set_bed(np.zeros((1, 1)))
    sim.set_state(np.array([[h]]), np.array([[u]]), np.array([[v]]))
    sim.set_manning(n)
    sim.update_dudt(0, 0.0, dt)
    with_m = sim.array("dudt")[0], sim.array("dvdt")[0]
    sim2 = O.Sim(1, 1, 10.0, 10.0)
    sim2.set_bed(np.zeros((1, 1)))
    sim2.set_state(np.array([[h]]), np.array([[u]]), np.array([[v]]))
    sim2.update_dudt(0, 0.0, dt)
    without = sim2.array("dudt")[0], sim2.array("dvdt")[0]
    return with_m[0] - without[0], with_m[1] - without[1], sim


def test_manning_closed_form():
    h, u, v, n = 0.5, 1.0, -0.4, 0.03
    du, dv, sim = _single(h, u, v, n)
    inv_h = h / (h * h + 1e-3)
    sf = n * n * math.sqrt(u * u + v * v) * inv_h ** (4.0 / 3.0)
    assert du == pytest.approx(-G * sf * u, rel=1e-13)
    assert dv == pytest.approx(-G * sf * v, rel=1e-13)
    assert sim.array("friction_slope")[0] == pytest.approx(sf, rel=1e-14)


def test_manning_overshoot_cap_and_dry_cut():
    du, dv, _ = _single(0.01, u=2.0, v=0.0, n=5.0, dt=2.0)
    assert du == pytest.approx(-2.0 / 2.0, rel=1e-15)       # capped to -u/dt
    du, dv, _ = _single(1e-7, u=2.0, v=1.0)
    assert du == 0.0 and dv == 0.0                          # h <= 1e-6: no friction


def test_pow43_and_exp_accuracy():
    L = O.lib()
    import ctypes as C
    L.or_pow43.restype = C.c_double; L.or_pow43.argtypes = [C.c_double]
    L.or_exp_neg.restype = C.c_double; L.or_exp_neg.argtypes = [C.c_double]
    for x in np.exp(np.linspace(math.log(1e-3), math.log(16.0), 200)):
        want = x ** (4.0 / 3.0)
        assert abs(L.or_pow43(x) - want) <= 8 * math.ulp(want)
    for x in np.linspace(-60.0, 0.0, 300):
        want = math.exp(x)
        assert abs(L.or_exp_neg(x) - want) <= 3 * math.ulp(want)
    assert L.or_exp_neg(-800.0) == 0.0


def _rain_cell(h, q0, q1, flags=0, t=10.0, dt=0.5, step=60.0):
    sim = O.Sim(1, 1, 10.0, 10.0, "heun2", flags)
    sim.set_bed(np.zeros((1, 1)))
    sim.set_state(np.array([[h]]))
    sim.set_time(0.0, 600.0, step, 1.0)
    sim.add_discharge([0.0, step, 1e6], [q0, q1, q1])
    sim.start_new_step(0.0)
    sim.update_dudt(0, t, dt)
    return sim.array("dhdt")[0]


def test_discharge_closed_form():
    q = 2e-5
    assert _rain_cell(1.0, q, q) == pytest.approx(q, rel=1e-15)
    # linear ramp over the step, averaged over [t, t + dt]
    q0, q1, t, dt, L = 1e-5, 3e-5, 10.0, 0.5, 60.0
    want = q0 + (q1 - q0) / L * (t + 0.5 * dt)
    assert _rain_cell(1.0, q0, q1, t=t, dt=dt) == pytest.approx(want, rel=1e-14)


def test_discharge_dry_guard_quirk_and_e1():
    # DischargeBoundarySourceTerm.cpp:62: h - dh*dt <= 0 -> dh = -h/dt: rain
    # never wets a dry cell in the reference
    assert _rain_cell(0.0, 2e-5, 2e-5) == 0.0
    assert _rain_cell(1e-6, 2e-5, 2e-5, dt=0.5) == pytest.approx(-1e-6 / 0.5)
    # E1: the guard only stops a negative rate from over-draining
    assert _rain_cell(0.0, 2e-5, 2e-5, flags=O.E1_RAIN_GUARD) == pytest.approx(2e-5)


def test_head_boundary_overwrites():
    sim = O.Sim(3, 1, 10.0, 10.0)
    sim.set_bed(np.zeros((1, 3)))
    sim.set_state(np.array([[2.0, 2.0, 2.0]]))
    sim.set_time(0.0, 600.0, 60.0, 1.0)
    sel = np.array([[1, 0, 0]], np.uint8)
    sim.add_head([0.0, 1e6], [0.5, 0.5], sel)
    sim.add_discharge([0.0, 1e6], [1e-3, 1e-3])
    sim.start_new_step(0.0)
    sim.update_dudt(0, 5.0, 0.25)
    dh = sim.array("dhdt")
    assert dh[0] == pytest.approx((0.5 - 2.0) / 0.25)   # overwritten, rain ignored
    assert dh[1] == pytest.approx(1e-3)


# ---------------------------------------------------------------- RK, controller, time
@pytest.mark.parametrize("name", sorted(O.TABLEAUX))
def test_tableau_order_conditions(name):
    n, m = O.tableau_matrix(name)
    b = m[n, 1:n + 1]
    assert abs(b.sum() - 1.0) < 1e-7
    for i in range(1, n):                     # c_i = sum_j a_ij
        assert abs(m[i, 0] - m[i, 1:i + 1].sum()) < 1e-7
    if n >= 2:                                # second order: sum b_i c_i = 1/2
        c = m[:n, 0]
        assert abs((b * c).sum() - 0.5) < 1e-7 or name in ("euler",)


def test_heun2_constant_rain_is_exact_and_time_linear_bias():
    q0, q1, L, dt = 1e-4, 3e-4, 60.0, 2.0
    sim = O.Sim(1, 1, 10.0, 10.0)
    sim.set_bed(np.zeros((1, 1)))
    sim.set_state(np.array([[1.0]]))
    sim.set_time(0.0, 600.0, L, 1.0)
    sim.add_discharge([0.0, L, 1e6], [q0, q1, q1])
    sim.start_new_step(0.0)
    t = 6.0
    sim.do_timestep(t, dt)
    h = sim.get_state(2)[0][0, 0]
    s = (q1 - q0) / L
    # stage 1 at t, stage 2 at t + dt, each averaging over its own [t_s, t_s + dt]
    want = 1.0 + dt * (q0 + s * (t + 0.5 * dt) + q0 + s * (t + 1.5 * dt)) / 2.0
    assert h == pytest.approx(want, rel=1e-15)


@pytest.mark.parametrize("dt,repeat,factor", [
    (50.0, True, None),     # cn > 5: dt/5
    (3.0, True, "cn"),      # 1 < cn <= 5: dt / (1.1 cn)
    (1.5, False, 1.0),      # 0.9 < cn <= 1: keep
    (0.5, False, "inv"),    # cn <= 0.9: dt / cn
])
def test_controller_table(dt, repeat, factor):
    sim = _lake(nx=4, ny=4, h=1.0, dx=10.0, dy=10.0)
    cn_expected = dt * 2.0 * math.sqrt(G * 1.0) / 10.0
    new_dt, rep = sim.do_timestep(0.0, dt)
    assert sim.control_number(2, dt) == pytest.approx(cn_expected, rel=1e-15)
    assert rep == repeat
    if factor is None:
        assert new_dt == dt / 5.0
    elif factor == "cn":
        assert new_dt == pytest.approx(dt / (cn_expected * 1.1), rel=1e-15)
    elif factor == 1.0:
        assert new_dt == dt
    else:
        assert new_dt == pytest.approx(dt / cn_expected, rel=1e-15)


def test_step_lands_exactly_and_counts():
    sim = _lake(nx=4, ny=4, h=1.0, dx=10.0, dy=10.0)
    sim.enable_trace(1000)
    sim.start_new_step(0.0)
    sim.step()
    tr = sim.trace()
    acc = tr[tr[:, 3] == 0]
    assert acc[-1, 0] + acc[-1, 1] == 60.0            # last accepted timestep ends on the step
    assert np.all(acc[:, 0] + acc[:, 1] <= 60.0)
    assert sim.stats()[0] == len(acc)


@pytest.mark.parametrize("end,n", [(3600.0, 60), (3601.0, 61), (3600.0 + 1e-7, 60), (30.0, 1)])
def test_num_steps(end, n):
    sim = _lake()
    sim.set_time(0.0, end, 60.0, 1.0)
    assert sim.num_steps() == n


def test_volume_of_constant_field():
    sim = _lake(nx=12, ny=9, h=0.75, dx=3.0, dy=2.0)
    assert sim.volume(0) == 0.75 * 12 * 9 * 3.0 * 2.0


# ---------------------------------------------------------------- physics extensions
def test_e8_dry_cells_carry_no_velocity_and_e9_caps():
    nx, ny = 6, 4
    sim = O.Sim(nx, ny, 1.0, 1.0, "heun2", O.E8_DRY_VELOCITY | O.E9_VELOCITY_CAP)
    sim.set_bed(np.zeros((ny, nx)))
    h = np.full((ny, nx), 1.0)
    h[:, :2] = 5e-4                        # below the 1 mm dry depth
    u = np.full((ny, nx), 30.0)            # above a 10 m/s cap
    sim.set_state(h, u, -u)
    sim.set_dry_depth(1e-3)
    sim.set_velocity_cap(10.0)
    sim.update_dudt(0, 0.0, 1e-6)
    sim.update_y(1, 1e-6)
    h1, u1, v1 = sim.get_state(1)
    assert np.all(u1[h1 <= 1e-3] == 0.0) and np.all(v1[h1 <= 1e-3] == 0.0)
    assert np.all(np.abs(u1) <= 10.0) and np.all(np.abs(v1) <= 10.0)
    assert np.any(u1 == 10.0)


# ------------------------------------------------ further sources/boundaries
def _flat_sim(n=5, h=1.0, u=0.0, v=0.0, dx=2.0, flags=0):
    sim = O.Sim(n, n, dx, dx, "heun2", flags)
    sim.set_bed(np.zeros((n, n)))
    sim.set_state(np.full((n, n), h), np.full((n, n), u), np.full((n, n), v))
    sim.set_time(0.0, 60.0, 60.0, 1.0)
    return sim


def test_energy_loss_closed_form_cap_and_dry_cut():
    # uniform flow on a flat bed: interior flux divergence and pressure are
    # exactly zero, so the interior dU/dt is the energy-loss term alone
    u, v, fx, fy, dt = 0.5, -0.3, 0.02, 0.04, 0.25
    sim = _flat_sim(u=u, v=v)
    sim.set_energy_loss(fx, fy)
    sim.start_new_step(0.0)
    sim.update_dudt(0, 0.0, dt)
    hyp = math.sqrt(u * u + v * v)
    assert sim.array("dudt").reshape(5, 5)[2, 2] == -0.5 * u * fx * hyp
    assert sim.array("dvdt").reshape(5, 5)[2, 2] == -0.5 * v * fy * hyp
    # overshoot: the loss would reverse u within dt -> -u/dt
    sim = _flat_sim(u=u, v=v)
    sim.set_energy_loss(1e4, 0.0)
    sim.start_new_step(0.0)
    sim.update_dudt(0, 0.0, dt)
    assert sim.array("dudt").reshape(5, 5)[2, 2] == -u / dt
    # no loss in dry cells (h <= 1e-6)
    sim = _flat_sim(h=5e-7, u=u, v=v)
    sim.set_energy_loss(fx, fy)
    sim.start_new_step(0.0)
    sim.update_dudt(0, 0.0, dt)
    assert sim.array("dudt").reshape(5, 5)[2, 2] == 0.0


@pytest.mark.parametrize("rate,cap,h,limit", [
    (1e-4, 0.1, 0.01, "rate"), (1.0, 0.1, 0.01, "depth"), (1.0, 0.002, 0.01, "capacity")])
def test_infiltration_limits_and_capacity_state(rate, cap, h, limit):
    dt = 0.5
    sim = _flat_sim(h=h)
    sim.set_infiltration(rate, cap)
    sim.start_new_step(0.0)
    sim.update_dudt(0, 0.0, dt)
    d = {"rate": rate * dt, "depth": h, "capacity": cap}[limit]
    assert sim.array("dhdt").reshape(5, 5)[2, 2] == -d / dt
    assert sim.array("infil_capacity").reshape(5, 5)[2, 2] == cap - d
    # the capacity is consumed again by every further application (every
    # stage, rejected timesteps included): apply once more to the same state
    sim.update_dudt(0, 0.0, dt)
    c1 = cap - d
    d2 = min(rate * dt, h, c1)
    assert sim.array("infil_capacity").reshape(5, 5)[2, 2] == c1 - d2
    assert sim.array("dhdt").reshape(5, 5)[2, 2] == -d2 / dt


def test_stage_boundary_is_head_relative_to_bed():
    n, dt, t = 5, 0.5, 10.0
    z = 2.0
    lev_t, lev_v = [0.0, 60.0, 1e6], [3.5, 4.1, 4.1]
    sel = np.zeros((n, n), np.uint8)
    sel[:, 0] = 1
    a = O.Sim(n, n, 2.0, 2.0, "heun2", 0)
    a.set_bed(np.full((n, n), z))
    a.set_state(np.full((n, n), 1.0))
    a.set_time(0.0, 60.0, 60.0, 1.0)
    a.add_stage(lev_t, lev_v, sel)
    a.start_new_step(0.0)
    a.update_dudt(0, t, dt)
    b = O.Sim(n, n, 2.0, 2.0, "heun2", 0)
    b.set_bed(np.full((n, n), z))
    b.set_state(np.full((n, n), 1.0))
    b.set_time(0.0, 60.0, 60.0, 1.0)
    b.add_head(lev_t, [x - z for x in lev_v], sel)
    b.start_new_step(0.0)
    b.update_dudt(0, t, dt)
    np.testing.assert_array_equal(a.array("dhdt"), b.array("dhdt"))
    # closed form at one boundary cell: target = mean of the level - bed at t, t+dt
    h0, h1 = 3.5 - z, 4.1 - z
    dh = (h1 - h0) / 60.0
    now = h0 + t * dh
    target = 0.5 * (now + (now + dt * dh))
    assert a.array("dhdt").reshape(n, n)[2, 0] == (target - 1.0) / dt
    assert a.array("dhdt").reshape(n, n)[2, 2] != (target - 1.0) / dt  # not selected


@pytest.mark.parametrize("flags", [0, O.E7_Y_ORIENT])
def test_eddy_viscosity_second_derivatives_of_quadratics(flags):
    n, dx = 12, 1.0
    a, b = 0.25, 0.125
    x = np.arange(n, dtype=np.float64)[None, :] * dx + np.zeros((n, 1))
    y = np.arange(n, dtype=np.float64)[:, None] * dx + np.zeros((1, n))
    sim = O.Sim(n, n, dx, dx, "heun2", flags)
    sim.set_bed(np.zeros((n, n)))
    sim.set_state(np.ones((n, n)), a * x * x, b * y * y)
    sim.set_time(0.0, 60.0, 60.0, 1.0)
    sim.set_eddy_viscosity(0.7)
    sim.start_new_step(0.0)
    sim.update_dudt(0, 0.0, 0.1)
    inner = (slice(3, n - 3), slice(3, n - 3))
    # Minmod3 of the Minmod3 slope of a x^2 is exactly 2a inside (both sign
    # flips of the reference's y operator cancel in d2/dy2)
    assert np.all(sim.array("d2udx2").reshape(n, n)[inner] == 2 * a)
    assert np.all(sim.array("d2vdy2").reshape(n, n)[inner] == 2 * b)
    assert np.all(sim.array("d2udy2").reshape(n, n)[inner] == 0.0)
    assert np.all(sim.array("d2vdx2").reshape(n, n)[inner] == 0.0)
    assert np.all(sim.array("eddy_mu").reshape(n, n) == 0.7)


@pytest.mark.parametrize("flags", [0, O.E7_Y_ORIENT])
def test_smagorinsky_coefficient_of_linear_fields(flags):
    n, dx, dy = 9, 2.0, 4.0
    al, be = 0.5, 0.25
    x = (np.arange(n, dtype=np.float64)[None, :] * dx) + np.zeros((n, 1))
    y = (np.arange(n, dtype=np.float64)[:, None] * dy) + np.zeros((1, n))
    sim = O.Sim(n, n, dx, dy, "heun2", flags)
    sim.set_bed(np.zeros((n, n)))
    sim.set_state(np.ones((n, n)), al * x, be * y)
    sim.set_time(0.0, 60.0, 60.0, 1.0)
    sim.set_eddy_viscosity(None, 0.05, 0.5)
    sim.start_new_step(0.0)
    sim.update_dudt(0, 0.0, 0.1)
    # du/dx = al, dv/dy = -+be (the reference's y slope is north->south
    # unless E7), cross terms 0: mu = c + s dx dy sqrt(al^2 + be^2)
    want = 0.05 + 0.5 * (dx * dy) * math.sqrt(al * al + be * be + 0.5 * (0.0 * 0.0))
    mu = sim.array("eddy_mu").reshape(n, n)
    assert np.all(mu[1:-1, 1:-1] == want)
    # linear fields: zero second derivatives, no eddy force
    assert np.all(sim.array("d2udx2").reshape(n, n)[1:-1, 1:-1] == 0.0)


# ---------------------------------------------------------------- E2 outflow edge
def _edge_sim(u, flags, nx=6, ny=4, h=0.8, v=0.0, dx=5.0, dy=4.0):
    sim = O.Sim(nx, ny, dx, dy, "heun2", flags)
    sim.set_bed(np.zeros((ny, nx)))
    sim.set_state(np.full((ny, nx), h), np.full((ny, nx), u), np.full((ny, nx), v))
    sim.set_time(0.0, 60.0, 60.0, 1.0)
    sim.start_new_step(0.0)
    sim.update_dudt(0, 0.0, 0.1)
    return sim


def test_e2_uniform_outflow_leaves_through_the_west_edge():
    """A uniform state flowing west (u < 0) on a flat bed: with E2 the x = 0
    faces pass the physical flux of that state (h u, u^2/2 in the reference's
    momentum form, v u), so the edge column keeps its depth; a closed wall
    (the reference) stops the water there and it piles up at -h u / dx."""
    h, u, v, dx = 0.8, -0.5, 0.1, 5.0
    nx, ny = 6, 4
    nxf = (nx + 1) * ny
    sim = _edge_sim(u, O.E2_WEST_OUTFLOW, h=h, v=v, dx=dx)
    west = lambda name: sim.array(name)[:nxf].reshape(ny, nx + 1)[:, 0]
    assert np.all(west("hflux") == h * u)
    np.testing.assert_allclose(west("uflux"), 0.5 * u * u, rtol=1e-14, atol=1e-15)
    assert np.all(west("vflux") == v * u)
    assert np.all(west("zflux") == 0.0)
    assert np.all(west("branch") == 1.0)
    # rows away from the north/south walls (whose v flux drains one wall row
    # and fills the other)
    dh = sim.array("dhdt").reshape(ny, nx)
    assert np.all(dh[1:-1, 0] == 0.0)
    # the east wall is closed with or without E2 (the flux only leaves west)
    east = sim.array("hflux")[:nxf].reshape(ny, nx + 1)[:, nx]
    assert np.all(east == 0.0)
    wall = _edge_sim(u, 0, h=h, v=v, dx=dx)
    dhw = wall.array("dhdt").reshape(ny, nx)
    np.testing.assert_allclose(dhw[1:-1, 0], -h * u / dx, rtol=1e-14)


def test_e2_ghost_carries_no_inflow_velocity():
    """Flow away from the x = 0 edge (u > 0): the E2 ghost takes the cell's
    depth and tangential velocity but zero normal velocity (min(u, 0)), so the
    face is the Rusanov average of (h, 0) and (h, u): mass h u / 2, normal
    momentum (u^2/2)/2 - (|u| + c) u / 2 with c = sqrt(g h).  A ghost that
    copied u (max instead of min) would pass h u."""
    h, u, dx = 0.8, 0.5, 5.0
    nx, ny = 6, 4
    nxf = (nx + 1) * ny
    sim = _edge_sim(u, O.E2_WEST_OUTFLOW, h=h, dx=dx)
    west = lambda name: sim.array(name)[:nxf].reshape(ny, nx + 1)[:, 0]
    c = math.sqrt(G * h)
    np.testing.assert_allclose(west("hflux"), 0.5 * h * u, rtol=1e-15)
    np.testing.assert_allclose(west("uflux"), 0.25 * u * u - 0.5 * (u + c) * u, rtol=1e-13)
    assert np.all(west("branch") == 1.0)


# ---------------------------------------------------------------- E4 storm
STORM = dict(peak=80.0 / 3.6e6, sigma=400.0, cx0=1000.0, cy0=700.0, vx=3.0, vy=-4.0)


def test_e4_storm_rate_closed_form():
    s = STORM
    # centre at t = 0, and the centre drifted by (vx t, vy t)
    assert O.storm_rate(s, s["cx0"], s["cy0"], 0.0) == s["peak"]
    t = 250.0
    assert O.storm_rate(s, s["cx0"] + s["vx"] * t, s["cy0"] + s["vy"] * t, t) == s["peak"]
    # one sigma along each axis, and along both (separable product)
    want1 = s["peak"] * math.exp(-0.5)
    assert O.storm_rate(s, s["cx0"] + s["sigma"], s["cy0"], 0.0) == pytest.approx(want1, rel=1e-15)
    assert O.storm_rate(s, s["cx0"], s["cy0"] - s["sigma"], 0.0) == pytest.approx(want1, rel=1e-15)
    assert O.storm_rate(s, s["cx0"] + s["sigma"], s["cy0"] + s["sigma"], 0.0) == \
        pytest.approx(s["peak"] * math.exp(-1.0), rel=1e-15)
    rng = np.random.default_rng(7)
    for _ in range(50):
        x, y, tt = rng.uniform(-500, 3000), rng.uniform(-500, 2500), rng.uniform(0, 900)
        dx, dy = x - (s["cx0"] + s["vx"] * tt), y - (s["cy0"] + s["vy"] * tt)
        want = s["peak"] * math.exp(-(dx * dx + dy * dy) / (2 * s["sigma"] ** 2))
        assert O.storm_rate(s, x, y, tt) == pytest.approx(want, rel=1e-13, abs=1e-300)
    # far from the centre the rate underflows to exactly zero
    assert O.storm_rate(s, s["cx0"] + 60 * s["sigma"], s["cy0"], 0.0) == 0.0


def test_e4_storm_feeds_the_discharge_term():
    """The storm is sampled at every cell centre at the outer step's two
    endpoint times and enters the discharge term as a linearly varying rate,
    averaged over [t_sub, t_sub + dt] (DischargeBoundarySourceTerm.cpp:43-67)."""
    nx, ny, dx, dy = 5, 3, 200.0, 300.0
    sim = O.Sim(nx, ny, dx, dy, "heun2", O.E4_STORM | O.E1_RAIN_GUARD)
    sim.set_bed(np.zeros((ny, nx)))
    sim.set_state(np.full((ny, nx), 1.0))
    sim.set_storm(STORM)
    sim.set_time(0.0, 600.0, 60.0, 1.0)
    t0, t_sub, dt, L = 120.0, 7.5, 0.5, 60.0
    sim.start_new_step(t0)
    sim.update_dudt(0, t_sub, dt)
    dh = sim.array("dhdt").reshape(ny, nx)
    s = STORM
    for y in range(ny):
        for x in range(nx):
            cx, cy = (x + 0.5) * dx, (y + 0.5) * dy

            def rate(t):
                ddx, ddy = cx - (s["cx0"] + s["vx"] * t), cy - (s["cy0"] + s["vy"] * t)
                return s["peak"] * math.exp(-(ddx * ddx + ddy * ddy) / (2 * s["sigma"] ** 2))
            q0, q1 = rate(t0), rate(t0 + L)
            want = q0 + (q1 - q0) / L * (t_sub + 0.5 * dt)
            assert dh[y, x] == pytest.approx(want, rel=1e-12), (x, y)


# ---------------------------------------------------------------- Manning blend
@pytest.mark.parametrize("h,t", [(0.05, 0.0), (0.1, 0.0), (0.15, 0.25), (0.2, 0.5),
                                 (0.25, 0.75), (0.3, 1.0), (0.9, 1.0)])
def test_manning_shallow_deep_blend(h, t):
    """n = mix(n_sh, n_dp, smoothstep(d_sh, d_dp, h)) with distinct n values
    (ManningRoughnessSourceTerm.cpp:57-60): n_sh below d_sh, n_dp above d_dp,
    the Hermite blend t^2 (3 - 2t) between; the friction uses that n."""
    n_sh, n_dp, d_sh, d_dp = 0.02, 0.08, 0.1, 0.3
    u, v = 0.6, -0.25
    sim = O.Sim(1, 1, 10.0, 10.0)
    sim.set_bed(np.zeros((1, 1)))
    sim.set_state(np.array([[h]]), np.array([[u]]), np.array([[v]]))
    sim.set_manning(n_sh, n_dp, d_sh, d_dp)
    sim.update_dudt(0, 0.0, 0.01)
    n_want = n_sh + (n_dp - n_sh) * (t * t * (3.0 - 2.0 * t))
    assert sim.array("mannings_n")[0] == pytest.approx(n_want, rel=1e-14, abs=1e-17)
    inv_h = h / (h * h + 1e-3)
    sf = n_want * n_want * math.sqrt(u * u + v * v) * inv_h ** (4.0 / 3.0)
    assert sim.array("friction_slope")[0] == pytest.approx(sf, rel=1e-13)
    # no other term acts on a single flat cell away from walls' momentum:
    # compare with the same cell without friction
    du, dv, _ = _single(h, u, v, n=n_want, dt=0.01)
    sim_du, sim_dv = sim.array("dudt")[0], sim.array("dvdt")[0]
    ref = O.Sim(1, 1, 10.0, 10.0)
    ref.set_bed(np.zeros((1, 1)))
    ref.set_state(np.array([[h]]), np.array([[u]]), np.array([[v]]))
    ref.update_dudt(0, 0.0, 0.01)
    assert sim_du - ref.array("dudt")[0] == pytest.approx(-G * sf * u, rel=1e-12)
    assert sim_dv - ref.array("dvdt")[0] == pytest.approx(-G * sf * v, rel=1e-12)


# ---------------------------------------------------------------- 40/60 landing
def _still_lake_dts(period_s, step=60.0, dt0=1.0):
    """A lake at rest whose control number is dt / period_s (dx = dy = 100 m,
    h chosen so that 2 sqrt(g h) / dx = 1 / period_s); returns the accepted
    timesteps of the first outer step."""
    dx = 100.0
    c = dx / (2.0 * period_s)
    h = c * c / G
    sim = O.Sim(3, 3, dx, dx)
    sim.set_bed(np.zeros((3, 3)))
    sim.set_state(np.full((3, 3), h))
    sim.set_time(0.0, step, step, dt0)
    sim.enable_trace(100)
    sim.solve()
    tr = sim.trace()
    return tr[tr[:, 3] == 0][:, 1], tr


def test_landing_splits_the_last_two_timesteps_40_60():
    """TemporalScheme.hpp:134-145.  After a 1 s first timestep the controller
    asks for ~40 s; 1 + 40 <= 60 but 1 + 2*40 > 60, so the next timestep is
    0.4 of the remaining 59 s and the last one lands on the step end."""
    dts, tr = _still_lake_dts(40.0)
    assert len(dts) == 3 and not np.any(tr[:, 3])
    assert dts[0] == 1.0
    assert dts[1] == (60.0 - 1.0) * 0.4
    assert dts[2] == 60.0 - (1.0 + (60.0 - 1.0) * 0.4)


def test_landing_goes_straight_to_the_step_end():
    """When the controller's timestep already overshoots the step end
    (1 + 80 > 60) the next timestep is the remainder, without the 40/60
    split."""
    dts, tr = _still_lake_dts(80.0)
    assert len(dts) == 2 and not np.any(tr[:, 3])
    assert dts[0] == 1.0 and dts[1] == 59.0
