"""HIP path against the CPU oracle on the paths and workloads beyond the
single-catchment plain method: coded-out (NaN-bed) cells, batch members,
served results, the wet/dry breach front (512^2 against the oracle, 4096^2
through properties), the 64-catchment ensemble, a serving stream drawn from
the request distribution, and the FP32 form on the storm catchment.

Bars (as in test_gpu_parity.py unless a test states its own):
  * one dU/dt evaluation: |k_gpu - k_oracle| <= 1e-12 * (1 + |k|) where the
    oracle's value is finite, the same NaN cells, identical branch codes;
  * runs: max |dh|, |d(hu)|, |d(hv)| <= 1e-9, wet masks (h > 1e-6) identical,
    FP64 volume relative 1e-10, identical accepted / rejected counts;
  * batches and servers: every member bit-identical to its alone run.
"""
import numpy as np
import pytest

from mfcm_amd import scenarios as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-configs]", *a, flush=True)


def _oracle(sc):
    from oracle import oracle as O
    return O.sim_from_scenario(sc)


def _coded_out_mask(ny, nx, seed):
    """Coded-out cells in every position the face logic distinguishes: the
    four corners, each edge, isolated interior cells, a 2x2 block, a
    horizontal run and a vertical pair, plus random ones."""
    rng = np.random.default_rng(seed)
    m = np.zeros((ny, nx), bool)
    for y, x in ((0, 0), (0, nx - 1), (ny - 1, 0), (ny - 1, nx - 1)):
        m[y, x] = True
    m[0, nx // 2] = m[ny - 1, nx // 3] = True
    m[ny // 2, 0] = m[ny // 3, nx - 1] = True
    if nx > 8 and ny > 8:
        m[ny // 2, nx // 2] = True
        m[2:4, 3:5] = True
        m[ny - 4, 2:6] = True
        m[ny // 4:ny // 4 + 2, nx - 3] = True
    k = max(1, nx * ny // 60)
    m[rng.integers(0, ny, k), rng.integers(0, nx, k)] = True
    return m


def _close(g, o, rel):
    """Finite values within rel * (1 + |o|), NaN exactly where the oracle has
    NaN."""
    np.testing.assert_array_equal(np.isnan(g), np.isnan(o))
    f = ~np.isnan(o)
    err = np.abs(g[f] - o[f]) / (1.0 + np.abs(o[f]))
    return float(err.max()) if err.size else 0.0


@pytest.mark.parametrize("nx,ny,dx,dy,flags,seed", [
    (64, 48, 5.0, 3.0, 0, 41),
    (61, 17, 2.0, 4.0, S.E7_Y_ORIENT, 42),       # ragged strips, power-of-two spacing
    (130, 33, 5.0, 3.0, S.E1_RAIN_GUARD | S.E8_DRY_VELOCITY, 43),
    (37, 29, 5.0, 5.0, S.E2_WEST_OUTFLOW | S.E7_Y_ORIENT, 44),
])
def test_dudt_coded_out_matches_oracle(gpu_lib, nx, ny, dx, dy, flags, seed):
    """FluxKernel.cpp:85-107 (a face next to a coded-out cell mirrors the
    wet side or carries nothing) and the NaN rules of Minmod3.hpp:60-81, on
    random wet/dry states over a bed with NaN cells at edges, corners, in
    clusters and alone."""
    from mfcm_amd.mfsv import Catchment
    rng = np.random.default_rng(seed)
    zb = rng.uniform(0.0, 2.0, (ny, nx))
    h = np.where(rng.uniform(size=(ny, nx)) < 0.6, rng.uniform(0.0, 1.5, (ny, nx)), 0.0)
    u = rng.normal(0, 0.7, (ny, nx)) * (h > 0)
    v = rng.normal(0, 0.7, (ny, nx)) * (h > 0)
    m = _coded_out_mask(ny, nx, seed)
    zb[m] = np.nan
    _say(f"dudt coded-out {nx}x{ny} flags={flags} seed={seed}: {int(m.sum())} NaN cells")
    sc = S.Scenario("nanbed", nx, ny, dx, dy, zb, h, u, v, 0.03, flags=flags, head_cols=1,
                    t_end=120.0, step_duration=30.0)
    sc.rain_t = np.array([0.0, 60.0, 90.0, 400.0])
    sc.rain_v = np.array([2e-5, 2e-5, 6e-6, 0.0])
    sim = _oracle(sc)
    sim.start_new_step(30.0)
    dt, t_sub = 0.37, 4.5
    sim.update_dudt(0, t_sub, dt)
    cat = Catchment(sc)
    cat.start_new_step(30.0)
    kh, ku, kv = cat.dudt(t_sub, dt, branches=True)
    for name, g in (("dhdt", kh), ("dudt", ku), ("dvdt", kv)):
        o = sim.array(name).reshape(ny, nx)
        # NaN rates sit on coded-out cells (their bed slope is NaN) and, in
        # the reference, on neighbours whose tangential bed slope is NaN
        # (two coded-out cells across them): same cells, same values
        err = _close(g, o, 1e-12)
        _say(f"  {name}: max rel err {err:.3g}, {int(np.sum(np.isnan(o)))} NaN "
             f"({int(np.sum(np.isnan(o[~m])))} outside coded-out cells)")
        assert err <= 1e-12, (name, err)
    br = sim.array("branch")
    nxf = (nx + 1) * ny
    np.testing.assert_array_equal(cat.download("branch_x").ravel(), br[:nxf])
    np.testing.assert_array_equal(cat.download("branch_y").ravel(), br[nxf:])
    cat.close()


def _isolated_mask(ny, nx):
    """Coded-out cells at the corners and edges and on a staggered interior
    lattice, no two within two cells of each other: no wet cell has a
    coded-out cell on both sides, so every slope stays finite."""
    m = np.zeros((ny, nx), bool)
    for y, x in ((0, 0), (0, nx - 1), (ny - 1, 0), (ny - 1, nx - 1), (0, nx // 2),
                 (ny - 1, nx // 3), (ny // 2, 0), (ny // 3, nx - 1)):
        m[y, x] = True
    for y in range(4, ny - 3, 7):
        for x in range(5 + (y % 3), nx - 3, 9):
            m[y, x] = True
    return m


def _coded_out_dambreak(isolated, t_end=60.0):
    sc = S.dam_break_bump(64, 48)
    sc.t_end = t_end
    m = _isolated_mask(sc.ny, sc.nx) if isolated else _coded_out_mask(sc.ny, sc.nx, 45)
    zb = sc.zb.copy()
    zb[m] = np.nan
    sc.zb = zb
    return sc, m


def _compare_run(cat, sim, tol=1e-9):
    hg, ug, vg = cat.state()
    ho, uo, vo = sim.get_state(0)
    np.testing.assert_array_equal(np.isnan(hg), np.isnan(ho))
    f = ~np.isnan(ho)
    assert np.abs(hg[f] - ho[f]).max() <= tol
    for a, b in ((hg * ug, ho * uo), (hg * vg, ho * vo)):
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        ff = ~np.isnan(b)
        assert np.abs(a[ff] - b[ff]).max() <= tol
    np.testing.assert_array_equal(hg > 1e-6, ho > 1e-6)
    return float(np.abs(hg[f] - ho[f]).max())


@pytest.mark.parametrize("isolated,graph", [(True, False), (True, True), (False, True)])
def test_run_coded_out_matches_oracle(gpu_lib, isolated, graph):
    """A dam break around a bump with coded-out cells at the edges, corners
    and inside the flow, 60 s, device controller (eager and graph).
    Isolated coded-out cells keep every depth finite; clustered ones (wet
    cells with a coded-out cell on both sides) make the reference's faces
    NaN there (a NaN tangential bed slope times the zero direction factor),
    and the NaN spreads through the stencil: the HIP path reproduces which
    cells turn NaN and every finite value."""
    from mfcm_amd.mfsv import Catchment
    sc, m = _coded_out_dambreak(isolated)
    _say(f"run coded-out dam break isolated={isolated} graph={graph}: {int(m.sum())} NaN cells")
    sim = _oracle(sc)
    sim.solve()
    cat = Catchment(sc)
    cat.run_begin()
    st = cat.run(graph=graph)
    err = _compare_run(cat, sim)
    acc, rej = sim.stats()
    assert (st["accepted"], st["rejected"]) == (acc, rej)
    # the reference lets NaN reach the depth of cells next to coded-out
    # ones: compare the volume of the finite cells (the NaN cells are equal)
    hg, ho = cat.state()[0], sim.get_state(0)[0]
    f = np.isfinite(ho)
    vg, vo = float(np.sum(hg[f])), float(np.sum(ho[f]))
    assert abs(vg - vo) <= 1e-10 * abs(vo)
    _say(f"  ok: {acc} accepted, {rej} rejected, max|dh| {err:.3g}, "
         f"{int(np.sum(~f))} NaN depths ({int(np.sum(~f & ~m))} outside coded-out cells)")
    if isolated:
        assert np.all(f)
    cat.close()


def test_batch_members_match_oracle(gpu_lib):
    """FP64 members of one batch (different shapes and spacings, one with a
    coded-out bed) each against the oracle run of that catchment alone
    (TemporalScheme.hpp:112-149 and RungeKutta.cpp:143-167 per member)."""
    from mfcm_amd import mfsv
    from mfcm_amd.batch import Batch
    scs = [S.ensemble_member(300 + k, side_lo=40, side_hi=110, t_end=120.0) for k in range(4)]
    zb = scs[2].zb.copy()
    zb[_coded_out_mask(scs[2].ny, scs[2].nx, 46)] = np.nan
    scs[2].zb = zb
    _say("batch members", [(s.nx, s.ny, round(s.dx, 2)) for s in scs])
    b = Batch(mfsv.F64)
    cats = []
    for sc in scs:
        c = mfsv.Catchment(sc, precision=mfsv.F64)
        c.run_begin()
        b.add(c)
        cats.append(c)
    assert b.run() == 0
    for sc, c in zip(scs, cats):
        sim = _oracle(sc)
        sim.solve()
        st = c.stats()
        assert st["done"] == 1
        assert (st["accepted"], st["rejected"]) == sim.stats()
        err = _compare_run(c, sim)
        vg, vo = c.volume(), sim.volume(0)
        assert abs(vg - vo) <= 1e-10 * max(abs(vo), 1e-30)
        _say(f"  member {sc.nx}x{sc.ny}: {st['accepted']}+{st['rejected']} timesteps, max|dh| {err:.3g}")
    b.close()
    for c in cats:
        c.close()


def test_served_results_match_oracle(gpu_lib):
    """The FP64 server's results (admitted under a budget that holds two of
    the four at a time) against the oracle run of each request."""
    from mfcm_amd import mfsv
    from mfcm_amd.serving import FloodServer, Request, footprint_bytes
    reqs = [Request(k, 700 + k, arrival=0.02 * k, side_lo=40, side_hi=100, horizon=90.0)
            for k in range(4)]
    scs = [r.scenario() for r in reqs]
    budget = sum(sorted(footprint_bytes(s.nx, s.ny, mfsv.F64) for s in scs)[-2:])
    srv = FloodServer(mfsv.F64, hbm_budget=budget, chunk=16)
    done = srv.serve(reqs, wall_limit=100.0, keep_fields=True)
    srv.close()
    assert len(done) == 4 and srv.backlog == 0
    for r in done:
        sc = r.scenario()
        sim = _oracle(sc)
        sim.solve()
        assert (r.result["accepted"], r.result["rejected"]) == sim.stats()
        ho = sim.get_state(0)[0]
        err = float(np.abs(r.result["h"] - ho).max())
        np.testing.assert_array_equal(r.result["h"] > 1e-6, ho > 1e-6)
        _say(f"  request {r.rid} {sc.nx}x{sc.ny}: max|dh| {err:.3g}")
        assert err <= 1e-9


def test_breach_512_matches_oracle(gpu_lib):
    """The wet/dry breach front (5 m of water behind a wall with a 50 m
    breach, dry bed downstream, dx 1 m) at 512^2 for 12 s, FP64 against the
    oracle."""
    from mfcm_amd.mfsv import Catchment
    sc = S.breach_front(seed=2, n=512, dx=1.0, t_end=12.0, wall_x=125, breach=50)
    sc.step_duration = 6.0
    _say(f"breach 512^2 12 s flags={sc.flags}")
    sim = _oracle(sc)
    sim.solve()
    cat = Catchment(sc)
    cat.run_begin()
    st = cat.run()
    err = _compare_run(cat, sim)
    acc, rej = sim.stats()
    assert (st["accepted"], st["rejected"]) == (acc, rej)
    vg, vo = cat.volume(), sim.volume(0)
    assert abs(vg - vo) <= 1e-10 * abs(vo)
    wet = int(np.sum(cat.state()[0][:, 126:] > 1e-6))
    _say(f"  ok: {acc} accepted, {rej} rejected, max|dh| {err:.3g}, {wet} wet cells downstream")
    assert wet > 0
    cat.close()


def test_breach_4096_properties(gpu_lib):
    """The breach front at its full 4096^2 size for 48 timesteps, FP64:
    the closed basin keeps its volume (FP64 mass balance within 1e-12), every
    value is finite, and the states are bit-identical for two tile heights."""
    from mfcm_amd.mfsv import Catchment
    sc = S.breach_front()
    _say(f"breach 4096^2 flags={sc.flags}")
    out = []
    for rows in (12, 36):
        cat = Catchment(sc)
        cat.set_tile_rows(rows)
        cat.run_begin()
        st = cat.run(48, graph=True)
        h, u, v = cat.state()
        assert np.all(np.isfinite(h)) and np.all(np.isfinite(u)) and np.all(np.isfinite(v))
        resid = st["volume"] - st["volume0"] - st["source_volume"]
        _say(f"  rows {rows}: {st['accepted']}+{st['rejected']} timesteps to t={st['time']:.3f} s, "
             f"volume {st['volume']:.6e}, mass residual {resid:.3g}")
        assert abs(resid) <= 1e-12 * st["volume0"]
        assert abs(st["volume"] - st["volume0"]) <= 1e-12 * st["volume0"]
        assert st["accepted"] + st["rejected"] == 48
        out.append((h, u, v, st["accepted"], st["rejected"]))
        cat.close()
    for x, y in zip(out[0][:3], out[1][:3]):
        np.testing.assert_array_equal(x, y)
    assert out[0][3:] == out[1][3:]


def test_ensemble_64_members_equal_alone(gpu_lib):
    """The 64-catchment ensemble (sides log-uniform in [128, 1024], FP32 in
    one launch per RK stage, hipGraph chunks) for 64 timesteps:
    every member finite, and a sample of members bit-identical to the same
    64 timesteps run alone."""
    from mfcm_amd import mfsv
    from mfcm_amd.autotune import model_rows_batch
    from mfcm_amd.batch import Batch
    scs = [S.ensemble_member(1000 + k) for k in range(64)]
    cells = sum(s.nx * s.ny for s in scs)
    rows = model_rows_batch([(s.nx, s.ny) for s in scs], mfsv.F32)
    _say(f"ensemble 64 members, {cells} cells, tile rows {rows}")
    b = Batch(mfsv.F32)
    cats = []
    for sc in scs:
        c = mfsv.Catchment(sc, precision=mfsv.F32)
        c.set_tile_rows(rows)
        c.run_begin()
        b.add(c)
        cats.append(c)
    b.run(64)
    for c in cats:
        st = c.stats()
        assert st["accepted"] + st["rejected"] == 64
        assert np.all(np.isfinite(c.download("h")))
    for k in (0, 17, 42, 63):
        alone = mfsv.Catchment(scs[k], precision=mfsv.F32)
        alone.set_tile_rows(rows)
        alone.run_begin()
        st = alone.run(64)
        sb = cats[k].stats()
        assert (st["accepted"], st["rejected"]) == (sb["accepted"], sb["rejected"])
        for x, y in zip(alone.state(), cats[k].state()):
            np.testing.assert_array_equal(x, y)
        _say(f"  member {k} {scs[k].nx}x{scs[k].ny}: equal alone ({st['accepted']} accepted)")
        alone.close()
    b.close()
    for c in cats:
        c.close()


def test_serving_stream_of_the_request_distribution(gpu_lib):
    """Requests drawn as the serving workload draws them (Poisson arrivals at
    20/s, sides log-uniform in [256, 2048], horizons uniform in [0.5, 3] h),
    FP32, admitted under a budget that holds only the two largest at once:
    every result bit-identical to the request run alone."""
    from mfcm_amd import mfsv
    from mfcm_amd.serving import FloodServer, footprint_bytes, poisson_stream, summarize
    reqs = poisson_stream(20.0, 0.3, seed=11, side_lo=256, side_hi=2048,
                          horizon_lo=1800.0, horizon_hi=10800.0)
    scs = [r.scenario() for r in reqs]
    _say("stream", [(s.nx, s.ny, round(s.dx, 1), round(r.horizon)) for s, r in zip(scs, reqs)])
    budget = sum(sorted(footprint_bytes(s.nx, s.ny, mfsv.F32) for s in scs)[-2:])
    srv = FloodServer(mfsv.F32, hbm_budget=budget, chunk=32)
    done = srv.serve(reqs, wall_limit=100.0, keep_fields=True)
    srv.close()
    assert len(done) == len(reqs) and srv.backlog == 0
    _say("summary", summarize(done, srv.elapsed))
    for r, sc in zip(sorted(done, key=lambda q: q.rid), scs):
        c = mfsv.Catchment(sc, precision=mfsv.F32)
        c.set_tile_rows(srv.tuner.rows_for(sc.nx, sc.ny, mfsv.F32))
        c.run_begin()
        st = c.run()
        assert (r.result["accepted"], r.result["rejected"]) == (st["accepted"], st["rejected"])
        np.testing.assert_array_equal(r.result["h"], c.download("h"))
        c.close()


def test_storm_catchment_256_f64_matches_oracle(gpu_lib):
    """The storm catchment's generator (hills and fBm on +300 m, Voronoi n,
    drifting storm, physics flags, CFL 0.5) at 256^2 for 120 s, FP64."""
    from mfcm_amd.mfsv import Catchment
    sc = S.storm_catchment(seed=1, n=256, dx=2.0, t_end=120.0)
    _say(f"storm catchment 256^2 120 s flags={sc.flags}")
    sim = _oracle(sc)
    sim.solve()
    cat = Catchment(sc)
    cat.run_begin()
    st = cat.run()
    err = _compare_run(cat, sim)
    assert (st["accepted"], st["rejected"]) == sim.stats()
    vg, vo = cat.volume(), sim.volume(0)
    assert abs(vg - vo) <= 1e-10 * abs(vo)
    _say(f"  ok: {st['accepted']}+{st['rejected']} timesteps, max|dh| {err:.3g}")
    cat.close()


def test_storm_catchment_f32_close_to_oracle(gpu_lib):
    """The FP32 form on the storm catchment (128^2, 240 s, the bed given to
    both as FP32 above its minimum) against the FP64 oracle, measured against
    the FP64 solution's own sensitivity: the oracle rerun with the bed
    perturbed by uniform noise of +-1e-7 m (a twentieth of the FP32 spacing
    of this bed).  The thin rain-fed sheet flow amplifies such perturbations
    to centimetres in single cells; the FP32 form must stay within 3x of what
    the perturbed FP64 run shows (max and 99th percentile of |dh|, relative
    volume), within 30 mm / 1 mm / 1e-3 outright, and close its own FP64 mass
    balance to 1e-7 of the volume."""
    from mfcm_amd.mfsv import Catchment, F32
    sc = S.storm_catchment(seed=1, n=128, dx=2.0, t_end=240.0)
    zmin = float(sc.zb.min())
    sc.zb = zmin + (sc.zb - zmin).astype(np.float32).astype(np.float64)
    sc.n_manning = np.asarray(sc.n_manning).astype(np.float32).astype(np.float64)
    sim = _oracle(sc)
    sim.solve()
    ho, vo = sim.get_state(0)[0], sim.volume(0)
    sp = S.storm_catchment(seed=1, n=128, dx=2.0, t_end=240.0)
    sp.n_manning = sc.n_manning
    sp.zb = sc.zb + np.random.default_rng(0).uniform(-1e-7, 1e-7, sc.zb.shape)
    simp = _oracle(sp)
    simp.solve()
    hp, vp = simp.get_state(0)[0], simp.volume(0)
    cat = Catchment(sc, precision=F32)
    cat.run_begin()
    st = cat.run()
    hg = cat.state()[0]
    vg = cat.volume()

    def metrics(h):
        d = np.abs(h - ho)
        wet = (h > 1e-6) | (ho > 1e-6)
        return d.max(), np.percentile(d[wet], 99)
    m32, p32 = metrics(hg)
    mp, pp = metrics(hp)
    v32, vpr = abs(vg - vo) / vo, abs(vp - vo) / vo
    resid = abs(st["volume"] - st["volume0"] - st["source_volume"]) / st["volume"]
    _say(f"storm f32 128^2 240 s: max|dh| {m32:.3g} m (FP64 +-1e-7 m bed: {mp:.3g}), "
         f"p99 {p32:.3g} ({pp:.3g}), rel volume {v32:.3g} ({vpr:.3g}), "
         f"FP32 mass residual {resid:.3g}, timesteps {st['accepted']}+{st['rejected']} vs "
         f"{sim.stats()} (perturbed {simp.stats()})")
    assert np.all(np.isfinite(hg))
    assert m32 <= max(3 * mp, 1e-3) and m32 <= 3e-2
    assert p32 <= max(3 * pp, 1e-5) and p32 <= 1e-3
    assert v32 <= max(3 * vpr, 1e-6) and v32 <= 1e-3
    assert resid <= 1e-7
    cat.close()


@pytest.mark.parametrize("case", ["hillslope_256_1h", "breach_1024_120s", "storm_256_300s"])
def test_f32_mass_balance_closes(gpu_lib, case):
    """The FP32 form's compensated depth keeps FP32 rounding out of its own
    FP64 mass balance (volume - volume0 - integrated sources, relative to the
    volume) over long runs: both Heun-2 stages carry what their FP32 depths
    round away in the low-order word, so the bound holds over thousands of
    timesteps (a form that dropped the first stage's rounding closed to
    5.5e-7 on the hillslope and 1e-7 on a breach front), and the timestep counts
    stay within 50 % of the FP64 run's (the thin rain-fed sheet flow of the
    hillslope takes more FP32 timesteps: films at the velocity cap set the
    CFL limit)."""
    from mfcm_amd.mfsv import Catchment, F32, F64
    if case == "hillslope_256_1h":
        sc, bound = S.hillslope_rain(seed=0, n=256, dx=5.0, physics=True, t_end=3600.0), 5e-8
    elif case == "breach_1024_120s":
        sc, bound = S.breach_front(seed=2, n=1024, dx=1.0, t_end=120.0, wall_x=250, breach=50), 1e-10
    else:
        sc, bound = S.storm_catchment(seed=1, n=256, dx=2.0, t_end=300.0), 5e-8
    res = {}
    for name, prec in (("f64", F64), ("f32", F32)):
        cat = Catchment(sc, precision=prec)
        cat.run_begin()
        st = cat.run()
        res[name] = st
        h = cat.state()[0]
        assert np.all(np.isfinite(h))
        cat.close()
    s32, s64 = res["f32"], res["f64"]
    resid = abs(s32["volume"] - s32["volume0"] - s32["source_volume"]) / abs(s32["volume"])
    r64 = abs(s64["volume"] - s64["volume0"] - s64["source_volume"]) / abs(s64["volume"])
    _say(f"{case}: FP32 mass residual {resid:.3g} (FP64 {r64:.3g}), timesteps FP32 "
         f"{s32['accepted']}+{s32['rejected']}, FP64 {s64['accepted']}+{s64['rejected']}")
    assert resid <= bound
    assert r64 <= 1e-12
    assert s32["time"] == s64["time"]
    assert s32["accepted"] <= 1.5 * s64["accepted"] and s64["accepted"] <= 1.5 * s32["accepted"]
